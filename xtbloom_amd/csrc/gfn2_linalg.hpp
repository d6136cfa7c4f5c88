// Dense FP64 linear algebra for the ragged batch: tiled GEMM over a tile
// queue, per-system Cholesky/inverse and the symmetric eigensolver
// (Householder tridiagonalization + implicit QL + back-transformation).
#pragma once

#include <type_traits>

#include "gfn2_device.hpp"
#include "gfn2_occupation.hpp"

namespace xtb {

// One output tile of one system's GEMM; the queue is built on the host from the
// ragged sizes so that one launch covers every system.
struct TileTask {
  int sys;
  int ti;  // tile row
  int tj;  // tile column
};

constexpr int kTile = 64;
constexpr int kTileK = 16;

// C[sys] = sum_k opA(A)[i][k] * scale[k] * opB(B)[k][j]   (row-major n x n operands)
//   TA: A is read transposed (A[k][i]); TB: B is read transposed (B[j][k]).
//   K per system: kdim[sys] if given, else n.  ncol (optional): tiles whose
//   columns all lie at or past ncol[sys] are skipped.
//   SYMM: only tiles with ti >= tj are queued; the kernel mirrors them.
//   TRI: the K range a tile needs when an operand is upper triangular (its
//   strict lower part stored as zero, as for X = L^-T), so the skipped
//   products are exact zeros:
//     1  A read transposed and upper triangular (X^T H):  k < i0 + 64
//     2  B upper triangular (T X):                         k < j0 + 64
//     3  A upper triangular (X Y):                         k >= i0
// FP64 matrix-core product of one 16 x 16 x 4 step, v_mfma_f64_16x16x4f64:
// lane l supplies A[l % 16][l / 16] and B[l / 16][l % 16] and accumulates
// C[l / 16 + 4 r][l % 16], r = 0..3 (layout checked on gfx950).
typedef double mfma_d4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ mfma_d4 mfma_16x16x4(double a, double b, mfma_d4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// Systems whose density product runs in split-precision FP32 (k_gemm_split32):
// those in their first SCC round or with a mixer residual above kSplitResidual.
constexpr double kSplitResidual = 1e-3;
__device__ __forceinline__ bool split_round(const SysState& z) {
  return z.iterations == 0 || z.residual_rms > kSplitResidual;
}

template <bool TA, bool TB, bool SYMM, int TRI = 0>
__global__ __launch_bounds__(256) void k_gemm(Topo t, const TileTask* tasks, const double* A,
                                              const double* Bm, double* C, const double* scale,
                                              const int* kdim, const int* ncol,
                                              const unsigned char* active,
                                              const SysState* split = nullptr) {
  const TileTask task = tasks[blockIdx.x];
  const int s = task.sys;
  if (active && !active[s]) return;
  if (split && split_round(split[s])) return;  // computed by k_gemm_split32
  if (ncol && task.tj * kTile >= ncol[s]) return;  // columns past ncol are not needed
  const int n = t.ao_off[s + 1] - t.ao_off[s];
  const int64_t mo = t.mat_off[s];
  const double* a = A + mo;
  const double* b = Bm + mo;
  double* c = C + mo;
  const double* sc = scale ? scale + t.ao_off[s] : nullptr;
  const int K = kdim ? kdim[s] : n;
  const int i0 = task.ti * kTile, j0 = task.tj * kTile;
  __shared__ double As[2][kTileK][kTile + 1];  // two stages: one written while the other is read
  __shared__ double Bs[2][kTileK][kTile + 1];
  // wave w owns the 32 x 32 quadrant (w / 2, w % 2): 2 x 2 blocks of 16 x 16
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int qi = 32 * (w >> 1), qj = 32 * (w & 1);
  const int l16 = lane & 15, lk = lane >> 4;
  mfma_d4 acc[2][2];
#pragma unroll
  for (int p = 0; p < 2; ++p)
#pragma unroll
    for (int q = 0; q < 2; ++q) acc[p][q] = mfma_d4{0.0, 0.0, 0.0, 0.0};
  int kbeg = 0, kend = K;
  if constexpr (TRI == 1) kend = min(K, i0 + kTile);
  if constexpr (TRI == 2) kend = min(K, j0 + kTile);
  if constexpr (TRI == 3) kbeg = i0;
  // Stage k0 of A and B is fetched into registers one step ahead: its global
  // loads are in flight while the matrix cores work on the previous stage.
  constexpr int kPer = kTile * kTileK / 256;  // elements of each operand per thread
  double pa[kPer], pb[kPer];
  auto fetch = [&](int k0) __attribute__((always_inline)) {
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
      const int idx = threadIdx.x + 256 * u;
      int kk, ii;
      if (TA) { ii = idx % kTile; kk = idx / kTile; } else { kk = idx % kTileK; ii = idx / kTileK; }
      const int gi = i0 + ii, gk = k0 + kk;
      double val = 0.0;
      if (gi < n && gk < K) {
        val = TA ? a[(int64_t)gk * n + gi] : a[(int64_t)gi * n + gk];
        if (sc) val *= sc[gk];
      }
      pa[u] = val;
    }
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
      const int idx = threadIdx.x + 256 * u;
      int kk, jj;
      if (TB) { kk = idx % kTileK; jj = idx / kTileK; } else { jj = idx % kTile; kk = idx / kTile; }
      const int gj = j0 + jj, gk = k0 + kk;
      pb[u] = (gj < n && gk < K) ? (TB ? b[(int64_t)gj * n + gk] : b[(int64_t)gk * n + gj]) : 0.0;
    }
  };
  // A tile rows i0..i0+63 -> As[buf][k][i], B tile -> Bs[buf][k][j]
  auto stash = [&](int buf) __attribute__((always_inline)) {
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
      const int idx = threadIdx.x + 256 * u;
      if (TA) As[buf][idx / kTile][idx % kTile] = pa[u]; else As[buf][idx % kTileK][idx / kTileK] = pa[u];
      if (TB) Bs[buf][idx % kTileK][idx / kTileK] = pb[u]; else Bs[buf][idx / kTile][idx % kTile] = pb[u];
    }
  };
  if (kbeg < kend) {
    fetch(kbeg);
    stash(0);
  }
  __syncthreads();
  int buf = 0;
  for (int k0 = kbeg; k0 < kend; k0 += kTileK, buf ^= 1) {
    const bool more = k0 + kTileK < kend;
    if (more) fetch(k0 + kTileK);
#pragma unroll
    for (int k4 = 0; k4 < kTileK; k4 += 4) {
      double af[2], bf[2];
#pragma unroll
      for (int p = 0; p < 2; ++p) af[p] = As[buf][k4 + lk][qi + 16 * p + l16];
#pragma unroll
      for (int q = 0; q < 2; ++q) bf[q] = Bs[buf][k4 + lk][qj + 16 * q + l16];
#pragma unroll
      for (int p = 0; p < 2; ++p)
#pragma unroll
        for (int q = 0; q < 2; ++q) acc[p][q] = mfma_16x16x4(af[p], bf[q], acc[p][q]);
    }
    // the other stage was last read before the previous barrier
    if (more) stash(buf ^ 1);
    __syncthreads();
  }
#pragma unroll
  for (int p = 0; p < 2; ++p)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int gi = i0 + qi + 16 * p + lk + 4 * r;
      if (gi >= n) continue;
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        const int gj = j0 + qj + 16 * q + l16;
        if (gj >= n) continue;
        c[(int64_t)gi * n + gj] = acc[p][q][r];
        if (SYMM && task.ti != task.tj) c[(int64_t)gj * n + gi] = acc[p][q][r];
      }
    }
}

// Split-precision FP32 matrix-core form of the symmetric density products
// P = C diag(f) C^T and W = C diag(f eps) C^T (lower tiles of the tile queue,
// K = kocc, mirrored).  Each FP64 operand x is split into x_hi = fp32(x) and
// x_lo = fp32(x - x_hi), and a_hi b_hi + a_hi b_lo + a_lo b_hi run on
// v_mfma_f32_16x16x4f32 with FP32 accumulation over one 16-deep K stage; the
// stage sums are added in FP64.  The dropped a_lo b_lo term is below 2^-48 of
// a product; the FP32 accumulation bounds the result near 1e-7 relative to
// the stage sums.  `only_early` (P): systems still in their first SCC round
// or with a mixer residual above kSplitResidual; null (W): every system.
// k_gemm skips the same systems through its `split` argument.

typedef float mfma_f4 __attribute__((ext_vector_type(4)));

__global__ __launch_bounds__(256) void k_gemm_split32(Topo t, const TileTask* tasks, const double* A,
                                                      double* C, const double* scale, const int* kdim,
                                                      const SysState* only_early,
                                                      const unsigned char* active) {
  const TileTask task = tasks[blockIdx.x];
  const int s = task.sys;
  if (active && !active[s]) return;
  if (only_early && !split_round(only_early[s])) return;
  const int n = t.ao_off[s + 1] - t.ao_off[s];
  const int64_t mo = t.mat_off[s];
  const double* a = A + mo;
  double* c = C + mo;
  const double* sc = scale + t.ao_off[s];
  const int K = kdim[s];
  const int i0 = task.ti * kTile, j0 = task.tj * kTile;
  // operands of one K stage: A = C[i][k] f_k (rows i0..), B^T = C[j][k] (rows j0..)
  __shared__ float Ah[kTileK][kTile + 1], Al[kTileK][kTile + 1];
  __shared__ float Bh[kTileK][kTile + 1], Bl[kTileK][kTile + 1];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int qi = 32 * (w >> 1), qj = 32 * (w & 1);
  const int l16 = lane & 15, lk = lane >> 4;
  double acc[2][2][4];
#pragma unroll
  for (int p = 0; p < 2; ++p)
#pragma unroll
    for (int q = 0; q < 2; ++q)
#pragma unroll
      for (int r = 0; r < 4; ++r) acc[p][q][r] = 0.0;
  constexpr int kPer = kTile * kTileK / 256;
  for (int k0 = 0; k0 < K; k0 += kTileK) {
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
      const int idx = threadIdx.x + 256 * u;
      const int kk = idx % kTileK, ii = idx / kTileK;
      const int gk = k0 + kk;
      double va = 0.0, vb = 0.0;
      if (gk < K) {
        if (i0 + ii < n) va = a[(int64_t)(i0 + ii) * n + gk] * sc[gk];
        if (j0 + ii < n) vb = a[(int64_t)(j0 + ii) * n + gk];
      }
      const float ha = (float)va, hb = (float)vb;
      Ah[kk][ii] = ha;
      Al[kk][ii] = (float)(va - (double)ha);
      Bh[kk][ii] = hb;
      Bl[kk][ii] = (float)(vb - (double)hb);
    }
    __syncthreads();
    mfma_f4 part[2][2];
#pragma unroll
    for (int p = 0; p < 2; ++p)
#pragma unroll
      for (int q = 0; q < 2; ++q) part[p][q] = mfma_f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int k4 = 0; k4 < kTileK; k4 += 4) {
      float ah[2], al[2], bh[2], bl[2];
#pragma unroll
      for (int p = 0; p < 2; ++p) {
        ah[p] = Ah[k4 + lk][qi + 16 * p + l16];
        al[p] = Al[k4 + lk][qi + 16 * p + l16];
      }
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        bh[q] = Bh[k4 + lk][qj + 16 * q + l16];
        bl[q] = Bl[k4 + lk][qj + 16 * q + l16];
      }
#pragma unroll
      for (int p = 0; p < 2; ++p)
#pragma unroll
        for (int q = 0; q < 2; ++q) {
          part[p][q] = __builtin_amdgcn_mfma_f32_16x16x4f32(al[p], bh[q], part[p][q], 0, 0, 0);
          part[p][q] = __builtin_amdgcn_mfma_f32_16x16x4f32(ah[p], bl[q], part[p][q], 0, 0, 0);
          part[p][q] = __builtin_amdgcn_mfma_f32_16x16x4f32(ah[p], bh[q], part[p][q], 0, 0, 0);
        }
    }
#pragma unroll
    for (int p = 0; p < 2; ++p)
#pragma unroll
      for (int q = 0; q < 2; ++q)
#pragma unroll
        for (int r = 0; r < 4; ++r) acc[p][q][r] += (double)part[p][q][r];
    __syncthreads();
  }
  // v_mfma_f32_16x16x4f32 leaves C[4 (l / 16) + r][l % 16] in lane l, item r
#pragma unroll
  for (int p = 0; p < 2; ++p)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int gi = i0 + qi + 16 * p + 4 * lk + r;
      if (gi >= n) continue;
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        const int gj = j0 + qj + 16 * q + l16;
        if (gj >= n) continue;
        c[(int64_t)gi * n + gj] = acc[p][q][r];
        if (task.ti != task.tj) c[(int64_t)gj * n + gi] = acc[p][q][r];
      }
    }
}

// Cholesky S = L L^T and X = L^{-T} with everything in global memory (any n);
// sets *bad on a non-positive pivot.
__device__ void cholesky_global(const double* Sm, double* L, double* Xm, int n, int& bad) {
  for (int64_t idx = threadIdx.x; idx < (int64_t)n * n; idx += blockDim.x) L[idx] = Sm[idx];
  __syncthreads();
  // right-looking column Cholesky
  for (int k = 0; k < n; ++k) {
    __shared__ double piv;
    if (threadIdx.x == 0) {
      double d = L[k + (int64_t)k * n];
      if (!(d > 0.0)) { bad = 1; d = 1.0; }
      piv = sqrt(d);
      L[k + (int64_t)k * n] = piv;
    }
    __syncthreads();
    for (int i = k + 1 + threadIdx.x; i < n; i += blockDim.x) L[i + (int64_t)k * n] /= piv;
    __syncthreads();
    // trailing update of the lower triangle: L[i][j] -= L[i][k] L[j][k], j in (k, i]
    const int m = n - k - 1;
    for (int64_t idx = threadIdx.x; idx < (int64_t)m * m; idx += blockDim.x) {
      const int jj = (int)(idx / m), ii = (int)(idx % m);
      if (ii < jj) continue;
      const int i = k + 1 + ii, j = k + 1 + jj;
      L[i + (int64_t)j * n] -= L[i + (int64_t)k * n] * L[j + (int64_t)k * n];
    }
    __syncthreads();
  }
  // inverse of L by columns: Linv e_j, thread per column j; X = Linv^T (row-major X[r][c] = Linv[c][r])
  // Linv is lower triangular; column j of Linv -> row j of X.
  for (int j = threadIdx.x; j < n; j += blockDim.x) {
    double* xr = Xm + (int64_t)j * n;  // row j of X holds column j of Linv
    for (int i = 0; i < j; ++i) xr[i] = 0.0;
    for (int i = j; i < n; ++i) {
      double v = (i == j) ? 1.0 : 0.0;
      for (int k = j; k < i; ++k) v -= L[i + (int64_t)k * n] * xr[k];
      xr[i] = v / L[i + (int64_t)i * n];
    }
  }
  __syncthreads();
}

// ---------------------------------------------------------------------------
// Symmetric eigensolver in three launches (all FP64):
//   k_tridiag         Householder reduction A = Q T Q^T, one 256-thread block per
//                     system; T's diagonal/off-diagonal go to d/e, the reflectors
//                     (H = I - 2 v v^T, |v| = 1) stay in A's strict lower triangle.
//   k_tridiag_eigen   eigenpairs of T, one block per system: T is split where an
//                     off-diagonal is negligible, every eigenvalue is found by
//                     Sturm-count bisection (one thread per eigenvalue), every
//                     eigenvector by inverse iteration with a partially pivoted
//                     tridiagonal LU (one thread per cluster of close eigenvalues,
//                     modified Gram-Schmidt inside a cluster).  Eigenvalues are
//                     ranked ascending (ties by index).
//   k_backtransform   Y = Q Z on column blocks staged in LDS, many blocks per
//                     system; column j of Z lands in column rank[j] of Y.
// Layouts: Z and Y are row-major (row i = component i, column j = vector j).
// ---------------------------------------------------------------------------
constexpr double kEps = 2.220446049250313e-16;
constexpr double kSafeMin = 2.2250738585072014e-308;
constexpr double kClusterTol = 1e-6;  // relative gap (to the block 1-norm) below which vectors are orthogonalized
// Clusters of more than kSerialCluster eigenvalues go to k_cluster_vectors
// (a near-dependent basis inflates |T| and can chain hundreds of eigenvalues
// into one cluster, whose Gram-Schmidt is quadratic in its size)
constexpr int kSerialCluster = 8;
// Hand-over of those clusters, per system: flag[s] set when it has one; for
// a cluster of indices j .. g1-1 (block-local eigenvalue order), end[ao + j]
// = g1 and lam / rank[ao + jj] hold each member's eigenvalue and Z column.
// k_cluster_vectors clears end[] entries after use.
struct ClusterOut {
  int* flag;
  int* end;
  double* lam;
  int* rank;
};
constexpr int kInverseIterations = 3;

// Householder reduction with the trailing matrix in global memory, one
// reflector at a time (any n).  The register kernel below keeps it as its
// general branch; the host sends systems above 256 AOs to the blocked
// reduction (k_tridiag_blocked), which reads the matrix a sixth as often.
__device__ __forceinline__ void tridiag_global(double* A, int n, double* d, double* e, double* v, double* pq,
                               double* red, double* bc) {
  const int tid = threadIdx.x, nt = blockDim.x;
  for (int k = 0; k + 2 < n; ++k) {
    const int m = n - k - 1;
    double ss = 0.0;
    for (int i = tid; i < m; i += nt) {
      const double x = A[(k + 1 + i) + (int64_t)k * n];
      v[i] = x;
      ss += x * x;
    }
    ss = block_sum(ss, red);
    if (tid == 0) {
      const double x0 = v[0];
      const double xn = sqrt(ss);
      d[k] = A[k + (int64_t)k * n];
      if (xn == 0.0 || ss - x0 * x0 == 0.0) {
        e[k] = x0;  // already reduced: no reflection
        bc[0] = 0.0;
      } else {
        const double alpha = x0 >= 0.0 ? -xn : xn;
        const double v0 = x0 - alpha;
        const double vv = ss - x0 * x0 + v0 * v0;
        e[k] = alpha;
        v[0] = v0;
        bc[0] = 1.0 / sqrt(vv);
      }
    }
    __syncthreads();
    const double inv = bc[0];
    if (inv == 0.0) {
      for (int i = tid; i < m; i += nt) A[(k + 1 + i) + (int64_t)k * n] = 0.0;
      __syncthreads();
      continue;
    }
    for (int i = tid; i < m; i += nt) {
      v[i] *= inv;
      A[(k + 1 + i) + (int64_t)k * n] = v[i];
    }
    __syncthreads();
    // p = A22 v, q = p - (v^T p) v
    double kp = 0.0;
    for (int i = tid; i < m; i += nt) {
      double p = 0.0;
      for (int j = 0; j < m; ++j) p += A[(k + 1 + i) + (int64_t)(k + 1 + j) * n] * v[j];
      pq[i] = p;
      kp += v[i] * p;
    }
    const double K = block_sum(kp, red);
    for (int i = tid; i < m; i += nt) pq[i] -= K * v[i];
    __syncthreads();
    // A22 -= 2 (v q^T + q v^T)
    for (int j = 0; j < m; ++j) {
      const double vj = v[j], qj = pq[j];
      for (int i = tid; i < m; i += nt)
        A[(k + 1 + i) + (int64_t)(k + 1 + j) * n] -= 2.0 * (v[i] * qj + pq[i] * vj);
    }
    __syncthreads();
  }
  if (tid == 0) {
    if (n >= 2) {
      d[n - 2] = A[(n - 2) + (int64_t)(n - 2) * n];
      e[n - 2] = A[(n - 1) + (int64_t)(n - 2) * n];
    }
    d[n - 1] = A[(n - 1) + (int64_t)(n - 1) * n];
    e[n - 1] = 0.0;
  }
}

// Householder reduction for n > 256 in panels of kTriPanel reflectors, the
// arrangement of LAPACK's dsytrd/dlatrd.  Within a panel the trailing matrix
// is not written: after reflectors 0 .. j-1 of the panel it equals
//   A_j = A_0 - V W^T - W V^T,
// V holding the panel's unit reflectors v_i (H_i = I - 2 v_i v_i^T, stored
// in A's columns below the subdiagonal as the register path stores them) and
// W the vectors w_i = 2 (p_i - (v_i^T p_i) v_i), p_i = A_i v_i (in the
// scratch matrix W, same layout).  Reflector j therefore reads column k of A_0
// corrected by V and W, the lower triangle of A_0's trailing block once (the
// symmetric product), and V and W twice; the trailing lower triangle is
// rewritten once per panel, on the matrix cores.  Only the lower triangle of
// A is read or written.  512 threads, n <= kMaxAO.
//
// Every phase issues all of its global loads before the first use (fully
// unrolled, out-of-range lanes clamped to valid addresses and their values
// selected away), so that a wavefront waits once per column pair rather than
// once per 64-row slot.
constexpr int kTriPanel = 16;
// LDS: `part` >= 8 kMaxAO doubles, `aux` >= kMaxAO + 4 kTriPanel doubles
constexpr int kTriAux = kMaxAO + 4 * kTriPanel;

// One unit of the symmetric product p = A v over an m x m trailing block
// (column-major at `blk`, leading dimension n, lower triangle): the columns
// c = cu .. cu + 2 kTriPairs - 1 (cu a multiple of 2 kTriPairs, so that all
// start in row slot t0 = cu / 64), rows 64 (t0 + u) + lane, u < NS.  Column
// c adds a_rc x_c to the row partials (registers, flushed to the wavefront's
// LDS row `prow` at the end) and its dot with x over the rows below the
// diagonal, plus the diagonal term, to pcol[c].  The next column pair is
// fetched while the current one is consumed.
//
// Rows past m read row m - 1 (a valid address) and meet x_r = 0, so they
// drop out of the dots without masks, and their row partials are never read;
// only slot 0, which holds the diagonal, is masked.  Column bases are
// wavefront-uniform and the row offsets 32-bit, so each load is one
// scalar-base instruction.
constexpr int kTriPairs = 8;

// Wavefront sums of two values at once: one exchange across the halves
// leaves column 0's partials in lanes 0..31 and column 1's in lanes 32..63,
// then five DPP steps sum each half (results in lanes 31 and 63).
__device__ __forceinline__ void wave_sum_pair(double d0, double d1, double& s0, double& s1) {
  const bool hi = (threadIdx.x & 32) != 0;
  double v = (hi ? d1 : d0) + __shfl_xor(hi ? d0 : d1, 32, 64);
  v += dpp_move<0xb1, 0xf>(v);   // quad_perm [1,0,3,2]
  v += dpp_move<0x4e, 0xf>(v);   // quad_perm [2,3,0,1]
  v += dpp_move<0x124, 0xf>(v);  // row_ror:4
  v += dpp_move<0x128, 0xf>(v);  // row_ror:8
  v += dpp_move<0x142, 0xa>(v);  // row_bcast:15 into rows 1 and 3
  const long long b = __double_as_longlong(v);
  s0 = __longlong_as_double(((long long)__builtin_amdgcn_readlane((int)(b >> 32), 31) << 32) |
                            (unsigned int)__builtin_amdgcn_readlane((int)b, 31));
  s1 = __longlong_as_double(((long long)__builtin_amdgcn_readlane((int)(b >> 32), 63) << 32) |
                            (unsigned int)__builtin_amdgcn_readlane((int)b, 63));
}

template <int NS>
__device__ __forceinline__ void symv_fetch(const double* __restrict__ blk, int64_t n, int m, int c,
                                           const unsigned (&boff)[NS], double (&f0)[NS], double (&f1)[NS]) {
  const char* p0 = reinterpret_cast<const char*>(blk + (int64_t)min(c, m - 1) * n);
  const char* p1 = reinterpret_cast<const char*>(blk + (int64_t)min(c + 1, m - 1) * n);
#pragma unroll
  for (int u = 0; u < NS; ++u) {
    f0[u] = *reinterpret_cast<const double*>(p0 + boff[u]);
    f1[u] = *reinterpret_cast<const double*>(p1 + boff[u]);
  }
}

template <int NS>
__device__ __forceinline__ void symv_consume(int c, int m, int r0, const double* __restrict__ x,
                                             const double (&xr)[NS], const double (&f0)[NS],
                                             const double (&f1)[NS], double (&acc)[NS],
                                             double* __restrict__ pcol, bool add) {
  const double xc0 = c < m ? x[c] : 0.0;
  const double xc1 = c + 1 < m ? x[c + 1] : 0.0;
  // slot 0: strictly lower rows, and the diagonal on the lane that holds it
  const double e0 = r0 > c ? f0[0] : 0.0;
  const double e1 = r0 > c + 1 ? f1[0] : 0.0;
  double d0 = fma(e0, xr[0], r0 == c ? f0[0] * xc0 : 0.0);
  double d1 = fma(e1, xr[0], r0 == c + 1 ? f1[0] * xc1 : 0.0);
  acc[0] = fma(e0, xc0, fma(e1, xc1, acc[0]));
#pragma unroll
  for (int u = 1; u < NS; ++u) {
    d0 = fma(f0[u], xr[u], d0);
    d1 = fma(f1[u], xr[u], d1);
    acc[u] = fma(f0[u], xc0, fma(f1[u], xc1, acc[u]));
  }
  wave_sum_pair(d0, d1, d0, d1);
  if ((threadIdx.x & 63) == 0) {
    if (c < m) pcol[c] = add ? pcol[c] + d0 : d0;
    if (c + 1 < m) pcol[c + 1] = add ? pcol[c + 1] + d1 : d1;
  }
}

// CH: the unit's rows start at slot tb instead of cu / 64 (row chunks of a
// trailing block above 16 slots); chunks after the first add their column
// dots to pcol, and their rows lie below every column of the unit.
template <int NS, bool CH = false>
__device__ __forceinline__ void symv_unit(const double* __restrict__ blk, int64_t n, int m, int cu, int tb,
                                          const double* __restrict__ x, double* __restrict__ prow,
                                          double* __restrict__ pcol) {
  const int lane = threadIdx.x & 63;
  const int t0 = CH ? tb : cu >> 6;
  const bool add = CH && tb != (cu >> 6);
  const int r0 = 64 * t0 + lane;  // this lane's row in slot 0
  unsigned boff[NS];              // byte offsets of this lane's rows in a column
  double xr[NS], acc[NS];
#pragma unroll
  for (int u = 0; u < NS; ++u) {
    const int rr = r0 + 64 * u;
    const int rc = min(rr, m - 1);
    boff[u] = (unsigned)rc * 8u;
    xr[u] = rr < m ? x[rc] : 0.0;
    acc[u] = 0.0;
  }
  // two column pairs per trip, in two buffers: pair q + 1 is in flight while
  // pair q is consumed, with no register copies between trips
  double a0[NS], a1[NS], b0[NS], b1[NS];
  symv_fetch<NS>(blk, n, m, cu, boff, a0, a1);
#pragma unroll 1
  for (int q = 0; q < kTriPairs; q += 2) {
    const int c = cu + 2 * q;
    symv_fetch<NS>(blk, n, m, c + 2, boff, b0, b1);
    symv_consume<NS>(c, m, r0, x, xr, a0, a1, acc, pcol, add);
    if (q + 2 < kTriPairs) symv_fetch<NS>(blk, n, m, c + 4, boff, a0, a1);
    symv_consume<NS>(c + 2, m, r0, x, xr, b0, b1, acc, pcol, add);
  }
#pragma unroll
  for (int u = 0; u < NS; ++u) prow[r0 + 64 * u] += acc[u];
}

template <int NS>
__device__ __forceinline__ void dots_pair(const double* __restrict__ col0, const double* __restrict__ col1,
                                          int m, const double* __restrict__ x, double* __restrict__ out0,
                                          double* __restrict__ out1, bool add = false) {
  const int lane = threadIdx.x & 63;
  double a0[NS], a1[NS];
#pragma unroll
  for (int u = 0; u < NS; ++u) {
    const int rr = min(64 * u + lane, m - 1);
    a0[u] = col0[rr];
    a1[u] = col1[rr];
  }
  double d0 = 0.0, d1 = 0.0;
#pragma unroll
  for (int u = 0; u < NS; ++u) {
    const int rr = 64 * u + lane;
    const double xv = rr < m ? x[min(rr, m - 1)] : 0.0;
    d0 = fma(a0[u], xv, d0);
    d1 = fma(a1[u], xv, d1);
  }
  wave_sum_pair(d0, d1, d0, d1);
  if (lane == 0) {
    *out0 = add ? *out0 + d0 : d0;
    *out1 = add ? *out1 + d1 : d1;
  }
}

#define XTB_SLOT_SWITCH(NSV, CALL)                                                   \
  switch (NSV) {                                                                     \
    case 1: { constexpr int NS = 1; CALL; } break;                                   \
    case 2: { constexpr int NS = 2; CALL; } break;                                   \
    case 3: { constexpr int NS = 3; CALL; } break;                                   \
    case 4: { constexpr int NS = 4; CALL; } break;                                   \
    case 5: { constexpr int NS = 5; CALL; } break;                                   \
    case 6: { constexpr int NS = 6; CALL; } break;                                   \
    case 7: { constexpr int NS = 7; CALL; } break;                                   \
    case 8: { constexpr int NS = 8; CALL; } break;                                   \
    case 9: { constexpr int NS = 9; CALL; } break;                                   \
    case 10: { constexpr int NS = 10; CALL; } break;                                 \
    case 11: { constexpr int NS = 11; CALL; } break;                                 \
    case 12: { constexpr int NS = 12; CALL; } break;                                 \
    case 13: { constexpr int NS = 13; CALL; } break;                                 \
    case 14: { constexpr int NS = 14; CALL; } break;                                 \
    case 15: { constexpr int NS = 15; CALL; } break;                                 \
    case 16: { constexpr int NS = 16; CALL; } break;                                 \
    default: break;                                                                  \
  }
static_assert(kMaxAO == 16 * 64, "the slot switch covers up to 16 slots of 64 rows");

// BIG (n > kMaxAO): v, pq, part and aux are global, with a row stride of
// `part` of n rounded up to 64; the symmetric product and the W^T v, V^T v
// dots run over row chunks of at most 16 slots.
template <bool BIG = false>
__device__ __forceinline__ void tridiag_blocked(double* __restrict__ A, double* __restrict__ W, int n,
                                                double* d, double* e, double* v, double* pq,
                                                double* part, double* aux, double* red, double* bc) {
  const int tid = threadIdx.x, nt = blockDim.x;
  // the wavefront index is uniform: keep it (and the column indices derived
  // from it) in scalar registers
  const int lane = tid & 63, wv = __builtin_amdgcn_readfirstlane(tid >> 6), nw = nt >> 6;
  const int ldp = BIG ? (n + 63) & ~63 : kMaxAO;
  double* pcol = aux;                   // [ldp] column dots of the symmetric product
  double* rowk = aux + ldp;             // [2 kTriPanel] row k of V, of W
  double* yv = rowk + 2 * kTriPanel;    // [2 kTriPanel] W^T v, V^T v
  double* prow = part + wv * ldp;       // this wavefront's row partials
  auto at = [n](int r, int c) __attribute__((always_inline)) { return (int64_t)r + (int64_t)c * n; };

  for (int k0 = 0; k0 + 2 < n; k0 += kTriPanel) {
    const int jn = min(kTriPanel, n - 2 - k0);
    for (int j = 0; j < jn; ++j) {
      const int k = k0 + j, b = k + 1, m = n - b;
      if (tid < kTriPanel) {
        const bool on = tid < j;
        const int ct = min(k0 + tid, n - 1);
        const double vk = A[at(k, ct)], wk = W[at(k, ct)];
        rowk[tid] = on ? vk : 0.0;
        rowk[kTriPanel + tid] = on ? wk : 0.0;
      }
      __syncthreads();
      // column k of A_j, rows k .. n-1 (the panel's later columns are read and
      // weighted by zero: one wait for all loads)
      double ss = 0.0;
      for (int r = k + tid; r < n; r += nt) {
        double av[kTriPanel], wv_[kTriPanel];
#pragma unroll
        for (int i = 0; i < kTriPanel; ++i) {
          av[i] = A[at(r, min(k0 + i, n - 1))];
          wv_[i] = W[at(r, min(k0 + i, n - 1))];
        }
        double x = A[at(r, k)];
#pragma unroll
        for (int i = 0; i < kTriPanel; ++i) {
          const bool on = i < j;
          x -= (on ? av[i] : 0.0) * rowk[kTriPanel + i] + (on ? wv_[i] : 0.0) * rowk[i];
        }
        if (r == k) {
          d[k] = x;
        } else {
          v[r - b] = x;
          ss += x * x;
        }
      }
      ss = block_sum(ss, red);
      if (tid == 0) {
        const double x0 = v[0];
        const double xn = sqrt(ss);
        if (xn == 0.0 || ss - x0 * x0 == 0.0) {
          e[k] = x0;  // already reduced: no reflection
          bc[0] = 0.0;
        } else {
          const double alpha = x0 >= 0.0 ? -xn : xn;
          const double v0 = x0 - alpha;
          const double vv = ss - x0 * x0 + v0 * v0;
          e[k] = alpha;
          v[0] = v0;
          bc[0] = 1.0 / sqrt(vv);
        }
      }
      __syncthreads();
      const double inv = bc[0];
      if (inv == 0.0) {
        for (int i = tid; i < m; i += nt) {
          A[at(b + i, k)] = 0.0;
          W[at(b + i, k)] = 0.0;
        }
        __syncthreads();
        continue;
      }
      const int tend = (m + 63) >> 6;
      for (int i = tid; i < m; i += nt) {
        const double vi = v[i] * inv;
        v[i] = vi;
        A[at(b + i, k)] = vi;
      }
      for (int i = lane; i < 64 * tend; i += 64) prow[i] = 0.0;
      __syncthreads();
      // p = A_0 v over the trailing block's lower triangle in units of
      // 2 kTriPairs columns, handed out in a snake over the wavefronts (the
      // units' work falls with their first column), then the 2j dots W^T v
      // and V^T v on the wavefront after the last unit's
      const int nunits = (m + 2 * kTriPairs - 1) / (2 * kTriPairs);
      for (int u0 = 0; u0 < nunits; u0 += nw) {
        const int u = u0 + (((u0 / nw) & 1) ? nw - 1 - wv : wv);
        if (u >= nunits) continue;
        const int cu = 2 * kTriPairs * u;
        const double* blk = A + at(b, b);
        if constexpr (BIG) {
          for (int tb = cu >> 6; tb < tend; tb += 16)
            XTB_SLOT_SWITCH(min(16, tend - tb), (symv_unit<NS, true>(blk, n, m, cu, tb, v, prow, pcol)))
        } else {
          XTB_SLOT_SWITCH(tend - (cu >> 6), (symv_unit<NS>(blk, n, m, cu, 0, v, prow, pcol)))
        }
      }
      for (int q = (wv + nw - nunits % nw) % nw; q < j; q += nw) {
        if constexpr (BIG) {
          for (int tb = 0; tb < tend; tb += 16)
            XTB_SLOT_SWITCH(min(16, tend - tb),
                            (dots_pair<NS>(W + at(b + 64 * tb, k0 + q), A + at(b + 64 * tb, k0 + q),
                                           m - 64 * tb, v + 64 * tb, yv + q, yv + kTriPanel + q, tb > 0)))
        } else {
          XTB_SLOT_SWITCH(tend, (dots_pair<NS>(W + at(b, k0 + q), A + at(b, k0 + q), m, v, yv + q,
                                                yv + kTriPanel + q)))
        }
      }
      __syncthreads();
      // p_j = A_j v = p - V (W^T v) - W (V^T v);  K = v^T p_j
      double kp = 0.0;
      for (int i = tid; i < m; i += nt) {
        double av[kTriPanel], wv_[kTriPanel];
#pragma unroll
        for (int q = 0; q < kTriPanel; ++q) {
          av[q] = A[at(b + i, min(k0 + q, n - 1))];
          wv_[q] = W[at(b + i, min(k0 + q, n - 1))];
        }
        double p = pcol[i];
        for (int w = 0; w < nw; ++w) p += part[w * ldp + i];
#pragma unroll
        for (int q = 0; q < kTriPanel; ++q) {
          const bool on = q < j;
          p -= (on ? av[q] : 0.0) * (on ? yv[q] : 0.0) + (on ? wv_[q] : 0.0) * (on ? yv[kTriPanel + q] : 0.0);
        }
        pq[i] = p;
        kp += v[i] * p;
      }
      const double K = block_sum(kp, red);
      for (int i = tid; i < m; i += nt) W[at(b + i, k)] = 2.0 * (pq[i] - K * v[i]);
      __syncthreads();
    }
    // trailing lower triangle, rows and columns >= c0:
    //   A -= V W^T + W V^T = [V W] [W V]^T
    // in 16 x 16 tiles on the matrix cores (K = 2 kTriPanel).  The product is
    // formed transposed, so that a lane's outputs run along a column of A
    // (row = l16, column = lk + 4 r) and the update reads and writes whole
    // column segments.
    const int c0 = k0 + jn, mt = n - c0;
    const int nb16 = (mt + 15) >> 4;
    const int ntiles = nb16 * (nb16 + 1) / 2;
    const int l16 = lane & 15, lk = lane >> 4;
    for (int tile = wv; tile < ntiles; tile += nw) {
      // tile -> (TI, TJ), TI >= TJ, row-major over the lower triangle
      int TI = (int)((sqrt(8.0 * tile + 1.0) - 1.0) * 0.5);
      while (TI * (TI + 1) / 2 > tile) --TI;
      while ((TI + 1) * (TI + 2) / 2 <= tile) ++TI;
      const int TJ = tile - TI * (TI + 1) / 2;
      const int rr = 16 * TI + l16, cc = 16 * TJ + l16;  // this lane's operand row and column
      const int rrc = c0 + min(rr, mt - 1), ccc = c0 + min(cc, mt - 1);
      double xa[2 * kTriPanel / 4], xb[2 * kTriPanel / 4];
#pragma unroll
      for (int q = 0; q < kTriPanel / 4; ++q) {
        const int i = 4 * q + lk;
        const int ci = k0 + min(i, jn - 1);
        // A operand: column side [W V] at row cc; B operand: row side [V W] at row rr
        xa[q] = W[at(ccc, ci)];
        xa[kTriPanel / 4 + q] = A[at(ccc, ci)];
        xb[q] = A[at(rrc, ci)];
        xb[kTriPanel / 4 + q] = W[at(rrc, ci)];
      }
      mfma_d4 acc = mfma_d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int q = 0; q < 2 * kTriPanel / 4; ++q) {
        const bool on = 4 * (q % (kTriPanel / 4)) + lk < jn;
        acc = mfma_16x16x4(on && cc < mt ? xa[q] : 0.0, on && rr < mt ? xb[q] : 0.0, acc);
      }
      // acc[r] = D[row = 16 TI + l16][col = 16 TJ + lk + 4 r]
      const int row = 16 * TI + l16;
      double old[4];
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int col = 16 * TJ + lk + 4 * r;
        old[r] = A[at(c0 + min(row, mt - 1), c0 + min(col, mt - 1))];
      }
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int col = 16 * TJ + lk + 4 * r;
        if (row < mt && col < mt && row >= col) A[at(c0 + row, c0 + col)] = old[r] - acc[r];
      }
    }
    __syncthreads();
  }
  if (tid == 0) {
    if (n >= 2) {
      d[n - 2] = A[(n - 2) + (int64_t)(n - 2) * n];
      e[n - 2] = A[(n - 1) + (int64_t)(n - 2) * n];
    }
    d[n - 1] = A[(n - 1) + (int64_t)(n - 1) * n];
    e[n - 1] = 0.0;
  }
}

// Householder reduction with the lower triangle held in registers, n <= 16 NR,
// 512 threads.  Thread (lr, lc) = (tid / 32, tid % 32) owns element
// (16 TI + lr, 32 TJ + lc) of every 16 x 32 tile that meets the lower triangle
// (TJ <= TI / 2), so the symmetric matvec and the rank-2 update of each step
// never touch memory.  Only the reflector (LDS, and its column of A for the
// back-transformation), the matvec partials (LDS) and d/e leave the registers.
// Stored form: strictly-lower elements as is, diagonal elements halved, the
// strictly-upper part of tiles crossing the diagonal as zero.  Then every slot
// contributes x v_j to p_i and x v_i to p_j, which is the full symmetric
// product, and the rank-2 update keeps the form with factor 2 / 1 / 0.
constexpr int tri_slots(int NR) { return NR == 0 ? 0 : tri_slots(NR - 1) + (NR - 1) / 2 + 1; }
// slot -> (row tile, column tile) of the register layout; folded at compile time
constexpr int tri_row(int slot) {
  int TI = 0;
  while (tri_slots(TI + 1) <= slot) ++TI;
  return TI;
}
constexpr int tri_col(int slot) { return slot - tri_slots(tri_row(slot)); }
// tile entirely strictly below the diagonal (every row index > every column index)
constexpr bool tri_strict(int TI, int TJ) { return 16 * TI > 32 * TJ + 31; }

// compile-time loop: f(std::integral_constant<int, I>) for I in [B, E)
template <int B, int E, typename F>
__device__ __forceinline__ void static_for(F&& f) {
  if constexpr (B < E) {
    f(std::integral_constant<int, B>{});
    static_for<B + 1, E>(f);
  }
}

constexpr int kTriPad = 257;  // row stride of the partial-sum buffers (bank spread)
constexpr int kTriTail = 64;    // trailing columns reduced by one wavefront
constexpr int kTriTailLd = 65;  // row stride of that trailing block in LDS
static_assert(kTriTail * kTriTailLd <= 32 * kTriPad, "the trailing block reuses the row partials");

template <int NR>
__device__ __forceinline__ void tridiag_registers(double* A, int n, double* d, double* e, double* v,
                                                  double* col, double* pq, double* rowpart,
                                                  double* colpart) {
  constexpr int NC = NR / 2;
  constexpr int NE = tri_slots(NR);
  const int tid = threadIdx.x, lr = tid >> 5, lc = tid & 31, lane = tid & 63;
  double a[NE];
  static_for<0, NE>([&](auto S) __attribute__((always_inline)) {
    constexpr int sl = decltype(S)::value;
    constexpr int TI = tri_row(sl), TJ = tri_col(sl);
    const int i = 16 * TI + lr, j = 32 * TJ + lc;
    // A is symmetric: lanes read row i along j contiguously
    const double x = (i < n && j < n) ? A[j + (int64_t)i * n] : 0.0;
    if constexpr (tri_strict(TI, TJ)) a[sl] = x;
    else a[sl] = i > j ? x : (i == j ? 0.5 * x : 0.0);
    // keep one row tile of loads in flight (address registers would spill the matrix)
    if constexpr (TJ == TI / 2) __builtin_amdgcn_sched_barrier(0);
  });
  // col holds column k of the reduced matrix (its owners publish column k+1
  // right after the update of step k); rows >= n stay zero.
  col[tid] = 0.0;
  __syncthreads();
  if (lc == 0) {
    static_for<0, NE>([&](auto S) __attribute__((always_inline)) {
      constexpr int sl = decltype(S)::value;
      constexpr int TI = tri_row(sl), TJ = tri_col(sl);
      if constexpr (TJ == 0) {
        const int i = 16 * TI + lr;
        if (i < n) col[i] = a[sl];
      }
    });
  }
  __syncthreads();
  int k = 0;
  for (; k + 2 < n && n - k > kTriTail; ++k) {
    // Reflector of column k, formed redundantly (and identically) by every
    // wavefront: v = (x - alpha e_{k+1}) / |x - alpha e_{k+1}|.
    double part = 0.0;
    for (int i = k + 1 + lane; i < n; i += 64) {
      const double x = col[i];
      part = fma(x, x, part);
    }
    const double ss = wave_sum(part);
    const double x0 = col[k + 1];
    const double xn = sqrt(ss);
    const bool noref = xn == 0.0 || ss - x0 * x0 == 0.0;  // already reduced
    const double alpha = x0 >= 0.0 ? -xn : xn;
    const double v0 = x0 - alpha;
    const double inv = noref ? 0.0 : 1.0 / sqrt(ss - x0 * x0 + v0 * v0);
    if (tid == 0) {
      d[k] = 2.0 * col[k];
      e[k] = noref ? x0 : alpha;
    }
    // no reflection (inv = 0) runs the same path with v = 0: p, q and the
    // update vanish and the stored reflector is zero
    const double vi = (tid > k && tid < n) ? inv * (tid == k + 1 ? v0 : col[tid]) : 0.0;
    v[tid] = vi;
    if (tid > k && tid < n) A[tid + (int64_t)k * n] = vi;
    __syncthreads();
    // p = A v (v = 0 outside the trailing block), column tile by column tile:
    // a column tile whose columns are all <= k is frozen (v = p = 0 there) and
    // skipped, which leaves about (n-k)^2/2 elements of work per step.  Lane
    // partials go to LDS, transposed so that thread i sums its row's 32 + 16
    // partials without cross-lane shuffles: rowpart[lc][i] from x v_j,
    // colpart[lr][j] from x v_i.
    double vr[NR], pr[NR];
#pragma unroll
    for (int TI = 0; TI < NR; ++TI) {
      vr[TI] = v[16 * TI + lr];
      pr[TI] = 0.0;
    }
    static_for<0, NC>([&](auto CJ) __attribute__((always_inline)) {
      constexpr int TJ = decltype(CJ)::value;
      if (32 * TJ + 31 > k) {
        const double vj = v[32 * TJ + lc];
        double pc = 0.0;
        static_for<2 * TJ, NR>([&](auto CI) __attribute__((always_inline)) {
          constexpr int TI = decltype(CI)::value;
          constexpr int sl = tri_slots(TI) + TJ;
          const double x = a[sl];
          pr[TI] = fma(x, vj, pr[TI]);
          pc = fma(x, vr[TI], pc);
        });
        colpart[lr * kTriPad + 32 * TJ + lc] = pc;
      }
    });
#pragma unroll
    for (int TI = 0; TI < NR; ++TI) rowpart[lc * kTriPad + 16 * TI + lr] = pr[TI];
    __syncthreads();
    // only partials of live tiles are read: row i > k gets rowpart from its own
    // row tile (written every step) and colpart from column tiles TJ with
    // 32 TJ + 31 >= i > k, all of them live
    double p = 0.0;
    if (tid > k && tid < n) {
      // eight independent partial sums: the loads overlap and the add chain is short
      double ps[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
      for (int r = 0; r < 32; r += 8) {
#pragma unroll
        for (int u = 0; u < 8; ++u) ps[u] += rowpart[(r + u) * kTriPad + tid];
      }
      for (int r = 0; r < 16; r += 8) {
#pragma unroll
        for (int u = 0; u < 8; ++u) ps[u] += colpart[(r + u) * kTriPad + tid];
      }
      p = ((ps[0] + ps[1]) + (ps[2] + ps[3])) + ((ps[4] + ps[5]) + (ps[6] + ps[7]));
    }
    pq[tid] = p;
    __syncthreads();
    // K = v.p, again per wavefront; with q = p - K v,
    // v_i q_j + q_i v_j = v_i p_j + (p_i - 2 K v_i) v_j.
    part = 0.0;
    for (int i = k + 1 + lane; i < n; i += 64) part = fma(v[i], pq[i], part);
    const double K2 = 2.0 * wave_sum(part);
    double rr[NR];
#pragma unroll
    for (int TI = 0; TI < NR; ++TI) rr[TI] = fma(-K2, vr[TI], pq[16 * TI + lr]);
    // A -= 2 (v q^T + q v^T) over the live column tiles
    static_for<0, NC>([&](auto CJ) __attribute__((always_inline)) {
      constexpr int TJ = decltype(CJ)::value;
      if (32 * TJ + 31 > k) {
        const int j = 32 * TJ + lc;
        const double vj = v[j], pj = pq[j];
        const double vjm = -2.0 * vj, pjm = -2.0 * pj;  // strict tiles: two FMAs per element
        static_for<2 * TJ, NR>([&](auto CI) __attribute__((always_inline)) {
          constexpr int TI = decltype(CI)::value;
          constexpr int sl = tri_slots(TI) + TJ;
          const int i = 16 * TI + lr;
          if constexpr (tri_strict(TI, TJ)) {
            a[sl] = fma(rr[TI], vjm, fma(vr[TI], pjm, a[sl]));
          } else {
            const double u = fma(vr[TI], pj, rr[TI] * vj);
            a[sl] -= (i > j ? 2.0 : (i == j ? 1.0 : 0.0)) * u;
          }
        });
      }
    });
    // publish column k+1 for the next step
    const int Tk1 = (k + 1) >> 5, ck1 = (k + 1) & 31;
    if (lc == ck1) {
      static_for<0, NE>([&](auto S) __attribute__((always_inline)) {
        constexpr int sl = decltype(S)::value;
        constexpr int TI = tri_row(sl), TJ = tri_col(sl);
        if (TJ == Tk1) {
          const int i = 16 * TI + lr;
          if (i < n) col[i] = a[sl];
        }
      });
    }
    __syncthreads();
  }
  if (k + 2 >= n) {  // n < 3: nothing was reduced
    static_for<0, NE>([&](auto S) __attribute__((always_inline)) {
      constexpr int sl = decltype(S)::value;
      constexpr int TI = tri_row(sl), TJ = tri_col(sl);
      const int i = 16 * TI + lr, j = 32 * TJ + lc;
      if (i == n - 2 && j == n - 2) d[n - 2] = 2.0 * a[sl];
      if (i == n - 1 && j == n - 2) e[n - 2] = a[sl];
      if (i == n - 1 && j == n - 1) d[n - 1] = 2.0 * a[sl];
    });
    if (tid == 0) e[n - 1] = 0.0;
    return;
  }
  // The last n - k0 <= kTriTail columns: one wavefront finishes them on the
  // trailing block, stored whole (symmetric, full diagonal) in LDS, without
  // block barriers.  Lane l owns row k + 1 + l of the step-k update.
  const int k0 = k;
  double* tb = rowpart;  // (n - k0) x kTriTailLd
  static_for<0, NE>([&](auto S) __attribute__((always_inline)) {
    constexpr int sl = decltype(S)::value;
    constexpr int TI = tri_row(sl), TJ = tri_col(sl);
    const int i = 16 * TI + lr, j = 32 * TJ + lc;
    if (i >= k0 && j >= k0 && i < n && j < n && i >= j) {
      const double x = a[sl];
      if (i > j) {
        tb[(i - k0) * kTriTailLd + (j - k0)] = x;
        tb[(j - k0) * kTriTailLd + (i - k0)] = x;
      } else {
        tb[(i - k0) * kTriTailLd + (i - k0)] = 2.0 * x;
      }
    }
  });
  __syncthreads();
  if (tid < 64) {
    const int l = tid;
    auto lds_fence = [] __attribute__((always_inline)) { asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory"); };
    for (; k + 2 < n; ++k) {
      const int m1 = n - k - 1;  // rows below the diagonal
      const int rl = k + 1 + l - k0, ck = k - k0;
      const double x = l < m1 ? tb[rl * kTriTailLd + ck] : 0.0;
      const double ss = wave_sum(x * x);
      const double x0 = tb[(ck + 1) * kTriTailLd + ck];
      const double xn = sqrt(ss);
      const bool noref = xn == 0.0 || ss - x0 * x0 == 0.0;  // already reduced
      const double alpha = x0 >= 0.0 ? -xn : xn;
      const double v0 = x0 - alpha;
      const double inv = noref ? 0.0 : 1.0 / sqrt(ss - x0 * x0 + v0 * v0);
      if (l == 0) {
        d[k] = tb[ck * kTriTailLd + ck];
        e[k] = noref ? x0 : alpha;
      }
      const double vl = l < m1 ? inv * (l == 0 ? v0 : x) : 0.0;
      if (l < m1) A[(k + 1 + l) + (int64_t)k * n] = vl;
      v[l] = vl;
      lds_fence();
      // p = A v over the trailing rows (v broadcast from LDS)
      double p = 0.0;
      if (l < m1) {
        const double* row = tb + rl * kTriTailLd + ck + 1;
        for (int j = 0; j < m1; ++j) p = fma(row[j], v[j], p);
      }
      // K = v.p; with q = p - K v,  A -= 2 (v q^T + q v^T) = 2 (v p^T + (p - 2 K v) v^T)
      const double K2 = 2.0 * wave_sum(vl * p);
      const double rq = fma(-K2, vl, p);
      pq[l] = p;
      lds_fence();
      if (l < m1) {
        double* row = tb + rl * kTriTailLd + ck + 1;
        for (int j = 0; j < m1; ++j) row[j] -= 2.0 * (vl * pq[j] + rq * v[j]);
      }
      lds_fence();
    }
    if (l == 0) {
      const int c2 = n - 2 - k0;
      d[n - 2] = tb[c2 * kTriTailLd + c2];
      e[n - 2] = tb[(c2 + 1) * kTriTailLd + c2];
      d[n - 1] = tb[(c2 + 1) * kTriTailLd + c2 + 1];
      e[n - 1] = 0.0;
    }
  }
  __syncthreads();
}

// The back-transformation applies the reflectors in compact WY blocks of
// kBackChunk:  H_{k_7} ... H_{k_0} = I - V T V^T,  V = [v_{k_0} .. v_{k_7}],
// k_u = n-3 - ch kBackChunk - u, T lower triangular with
//   T_mm = 2,  T_mi = -2 sum_{l=i}^{m-1} (v_m . v_l) T_li  (i < m),
// which follows from H_m (I - V T V^T) = I - [V v_m] [[T, 0], [-2 v_m^T V T, 2]] [V v_m]^T.
// T_mi of block ch sits at Tall[wy_offset + 64 ch + 8 m + i]; a system holds
// ceil((n-2)/8) blocks of 64 entries, at most 8 n + 40.
constexpr int kBackChunk = 8;  // reflectors per WY block
__device__ __forceinline__ int64_t wy_offset(const Topo& t, int s) {
  return (int64_t)kBackChunk * t.ao_off[s] + (int64_t)kBackChunk * kBackChunk * s;
}

// T factors of all WY blocks of one system from its stored reflectors (A's
// strict lower part, column k = reflector k), by the whole block: one thread
// per (block, pair) inner product over rows > k_i (where both reflectors live),
// then one thread per block runs the recurrence.  gsh holds 28 doubles per block.
// (Each caller takes its own instance: a second caller of one instance changes
// the register kernel's register allocation.)
template <int INSTANCE>
__device__ __forceinline__ void wy_factors(const double* A, int n, double* Ts, double* gsh) {
  constexpr int kPairs = kBackChunk * (kBackChunk - 1) / 2;  // 28
  const int tid = threadIdx.x, nt = blockDim.x;
  const int nref = n - 2, nch = (nref + kBackChunk - 1) / kBackChunk;
  for (int task = tid; task < nch * kPairs; task += nt) {
    const int ch = task / kPairs;
    int p = task - ch * kPairs, m = 1, i = 0;
    while (p > 0) {
      --p;
      if (++i == m) { ++m; i = 0; }
    }
    const int kb = n - 3 - ch * kBackChunk, km = kb - m, ki = kb - i;
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    if (km >= 0) {
      const double* vm = A + (int64_t)km * n;
      const double* vi = A + (int64_t)ki * n;
      int r = ki + 1;
      for (; r + 3 < n; r += 4) {
#pragma unroll
        for (int u = 0; u < 4; ++u) acc[u] = fma(vm[r + u], vi[r + u], acc[u]);
      }
      for (; r < n; ++r) acc[0] = fma(vm[r], vi[r], acc[0]);
    }
    gsh[task] = (acc[0] + acc[1]) + (acc[2] + acc[3]);
  }
  __syncthreads();
  for (int ch = tid; ch < nch; ch += nt) {
    const int kb = n - 3 - ch * kBackChunk;
    const double* g = gsh + ch * kPairs;  // (v_m . v_l) at m (m-1) / 2 + l
    double T[kBackChunk][kBackChunk];
#pragma unroll
    for (int m = 0; m < kBackChunk; ++m) {
#pragma unroll
      for (int i = 0; i < kBackChunk; ++i) T[m][i] = 0.0;
      if (kb - m < 0) continue;  // past the last reflector: zero row
      T[m][m] = 2.0;
#pragma unroll
      for (int i = 0; i < m; ++i) {
        double acc = 0.0;
#pragma unroll
        for (int l = i; l < m; ++l) acc = fma(g[m * (m - 1) / 2 + l], T[l][i], acc);
        T[m][i] = -2.0 * acc;
      }
    }
    double* To = Ts + ch * kBackChunk * kBackChunk;
#pragma unroll
    for (int m = 0; m < kBackChunk; ++m)
#pragma unroll
      for (int i = 0; i < kBackChunk; ++i) To[m * kBackChunk + i] = T[m][i];
  }
  __syncthreads();
}

// 512 threads per system (two waves per SIMD, 256 VGPRs per thread): the
// register reduction for n <= 256, the global-memory one otherwise (the host
// launches the systems of at most 256 AOs only)
__global__ __launch_bounds__(512) void k_tridiag(Topo t, const int* order, double* Aall, double* dall,
                                                 double* eall, double* Tall,
                                                 const unsigned char* active) {
  const int s = order[blockIdx.x];
  if (active && !active[s]) return;
  const int n = t.ao_off[s + 1] - t.ao_off[s];
  double* A = Aall + t.mat_off[s];  // column-major view of the symmetric matrix
  double* d = dall + t.ao_off[s];
  double* e = eall + t.ao_off[s];
  __shared__ double v[kMaxAO];
  __shared__ double pq[kMaxAO];
  __shared__ double col[512];  // published column of the register path
  __shared__ double rowpart[32 * kTriPad];
  __shared__ double colpart[16 * kTriPad];
  __shared__ double red[16];
  __shared__ double bc[2];
  if (n <= 64) tridiag_registers<4>(A, n, d, e, v, col, pq, rowpart, colpart);
  else if (n <= 128) tridiag_registers<8>(A, n, d, e, v, col, pq, rowpart, colpart);
  else if (n <= 192) tridiag_registers<12>(A, n, d, e, v, col, pq, rowpart, colpart);
  else if (n <= 256) tridiag_registers<16>(A, n, d, e, v, col, pq, rowpart, colpart);
  else tridiag_global(A, n, d, e, v, pq, red, bc);
  // WY factors for the register back-transformation (n <= 256); the stored
  // reflectors must be visible to the whole block first
  if (n <= 256 && n > 2) {
    __syncthreads();
    wy_factors<0>(A, n, Tall + wy_offset(t, s), rowpart);
  }
}

// n > 256: the blocked reduction, one 512-thread block per system (the host
// launches the large systems only); W is per-system n x n scratch.  BIG: the
// systems above kMaxAO AOs, with the work vectors in their global slice of `big`.
template <bool BIG>
__global__ __launch_bounds__(512) void k_tridiag_blocked(Topo t, const int* order, double* Aall,
                                                         double* dall, double* eall, double* Wall,
                                                         const unsigned char* active, double* big) {
  const int s = order[blockIdx.x];
  if (active && !active[s]) return;
  const int n = t.ao_off[s + 1] - t.ao_off[s];
  __shared__ double v_l[BIG ? 1 : kMaxAO];
  __shared__ double pq_l[BIG ? 1 : kMaxAO];
  __shared__ double part_l[BIG ? 1 : 8 * kMaxAO];
  __shared__ double aux_l[BIG ? 1 : kTriAux];
  static_assert(kTriPanel % 4 == 0, "panel in K steps of 4");
  __shared__ double red[16];
  __shared__ double bc[2];
  double *v = v_l, *pq = pq_l, *part = part_l, *aux = aux_l;
  if constexpr (BIG) {
    // 2 n + 8 ldp + ldp + 4 kTriPanel <= kBigPerAO n + kBigPerSys doubles
    const int ldp = (n + 63) & ~63;
    v = big + big_offset(t, s);
    pq = v + n;
    part = pq + n;
    aux = part + 8 * ldp;
  }
  tridiag_blocked<BIG>(Aall + t.mat_off[s], Wall + t.mat_off[s], n, dall + t.ao_off[s],
                       eall + t.ao_off[s], v, pq, part, aux, red, bc);
}

// Overlap Cholesky S = L L^T (L column-major into `L`) and X = L^{-T} (row-major,
// upper triangular, into `Xm`) with the lower triangle in registers (same
// 16 x 32 tile layout as tridiag_registers, 512 threads, n <= 16 NR):
//   factorization  per column k: the owners broadcast column k through LDS, the
//                  trailing lower triangle takes the rank-1 update l l^T / l_kk;
//   inverse        Gauss-Jordan on the identity: row k of L^{-1} is scaled by
//                  1/L_kk, then rows i > k lose L_ik times it.
template <int NR>
__device__ __forceinline__ void cholesky_registers(const double* Sm, double* L, double* Xm, int n,
                                                   double* lcol, double* xr, int& bad) {
  constexpr int NE = tri_slots(NR);
  const int tid = threadIdx.x, lr = tid >> 5, lc = tid & 31;
  double a[NE];
  static_for<0, NE>([&](auto S) __attribute__((always_inline)) {
    constexpr int sl = decltype(S)::value;
    constexpr int TI = tri_row(sl), TJ = tri_col(sl);
    const int i = 16 * TI + lr, j = 32 * TJ + lc;
    a[sl] = (i < n && j < n && i >= j) ? Sm[(int64_t)i * n + j] : 0.0;
    if constexpr (TJ == TI / 2) __builtin_amdgcn_sched_barrier(0);
  });
  // lcol is zero outside [k, n): the rank-1 update then needs no masks (column k
  // itself and the strict upper part of diagonal tiles take harmless updates
  // that are never read again)
  if (tid >= n) lcol[tid] = 0.0;
  for (int k = 0; k < n; ++k) {
    const int Tk = k >> 5, ck = k & 31;
    if (tid < k) lcol[tid] = 0.0;
    if (lc == ck) {
      static_for<0, NE>([&](auto S) __attribute__((always_inline)) {
        constexpr int sl = decltype(S)::value;
        constexpr int TI = tri_row(sl), TJ = tri_col(sl);
        if (TJ == Tk) {
          const int i = 16 * TI + lr;
          if (i >= k && i < n) lcol[i] = a[sl];
        }
      });
    }
    __syncthreads();
    double piv = lcol[k];
    if (!(piv > 0.0)) {
      if (tid == 0) bad = 1;
      piv = 1.0;
    }
    const double sq = sqrt(piv);
    if (tid >= k && tid < n) L[tid + (int64_t)k * n] = tid == k ? sq : lcol[tid] / sq;
    const double ip = 1.0 / piv;
    // column tiles entirely left of k are final (lcol is zero there): skipped
    static_for<0, NR / 2>([&](auto CJ) __attribute__((always_inline)) {
      constexpr int TJ = decltype(CJ)::value;
      if (32 * TJ + 31 >= k) {
        const double lj = lcol[32 * TJ + lc] * ip;
        static_for<2 * TJ, NR>([&](auto CI) __attribute__((always_inline)) {
          constexpr int TI = decltype(CI)::value;
          constexpr int sl = tri_slots(TI) + TJ;
          a[sl] -= lcol[16 * TI + lr] * lj;
        });
      }
    });
    __syncthreads();
  }
  // L^{-1} in the same registers
  static_for<0, NE>([&](auto S) __attribute__((always_inline)) {
    constexpr int sl = decltype(S)::value;
    constexpr int TI = tri_row(sl), TJ = tri_col(sl);
    const int i = 16 * TI + lr, j = 32 * TJ + lc;
    a[sl] = (i == j && i < n) ? 1.0 : 0.0;
  });
  // lcol holds column k of L below the diagonal (zero elsewhere), xr row k of
  // L^{-1} (zero right of the diagonal), so the update needs no masks either
  for (int k = 0; k < n; ++k) {
    if (tid < kMaxAO) {
      lcol[tid] = (tid > k && tid < n) ? L[tid + (int64_t)k * n] : 0.0;
      if (tid > k) xr[tid] = 0.0;
    }
    __syncthreads();
    const double il = 1.0 / L[k + (int64_t)k * n];
    const bool rowk = lr == (k & 15);
    static_for<0, NE>([&](auto S) __attribute__((always_inline)) {
      constexpr int sl = decltype(S)::value;
      constexpr int TI = tri_row(sl), TJ = tri_col(sl);
      const int j = 32 * TJ + lc;
      const bool own = rowk && TI == (k >> 4) && j <= k;
      const double nv = own ? a[sl] * il : a[sl];
      a[sl] = nv;  // select, not a branch: keeps the register array free of copies
      if (own) xr[j] = nv;
    });
    __syncthreads();
    // xr is zero right of column k: column tiles entirely right of k are skipped
    static_for<0, NR / 2>([&](auto CJ) __attribute__((always_inline)) {
      constexpr int TJ = decltype(CJ)::value;
      if (32 * TJ <= k) {
        const double xj = xr[32 * TJ + lc];
        static_for<2 * TJ, NR>([&](auto CI) __attribute__((always_inline)) {
          constexpr int TI = decltype(CI)::value;
          constexpr int sl = tri_slots(TI) + TJ;
          a[sl] -= lcol[16 * TI + lr] * xj;
        });
      }
    });
    __syncthreads();
  }
  // X[r][c] = Linv[c][r]: the registers hold Linv(i, j), i >= j; X's strict lower part is zero
  for (int64_t idx = tid; idx < (int64_t)n * n; idx += blockDim.x) {
    const int r = (int)(idx / n), c = (int)(idx - (int64_t)r * n);
    if (c < r) Xm[idx] = 0.0;
  }
  static_for<0, NE>([&](auto S) __attribute__((always_inline)) {
    constexpr int sl = decltype(S)::value;
    constexpr int TI = tri_row(sl), TJ = tri_col(sl);
    const int i = 16 * TI + lr, j = 32 * TJ + lc;
    if (i >= j && i < n) Xm[(int64_t)j * n + i] = a[sl];
    if constexpr (TJ == TI / 2) __builtin_amdgcn_sched_barrier(0);  // one row tile of stores in flight
  });
  __syncthreads();
}

// Per-system overlap factorization, 512 threads: X = L^{-T} with L in `work`
// (column-major).  status = 8 on a non-positive pivot; the conditioning gate
// follows in k_overlap_rcond.
__global__ __launch_bounds__(512) void k_cholesky_inverse(Topo t, const int* order, const double* S,
                                                          double* work, double* X, SysState* status) {
  const int s = order[blockIdx.x];
  const int n = t.ao_off[s + 1] - t.ao_off[s];
  const int64_t mo = t.mat_off[s];
  const double* Sm = S + mo;
  double* L = work + mo;
  double* Xm = X + mo;
  __shared__ double lcol[kMaxAO];
  __shared__ double xr[kMaxAO];
  __shared__ int bad;
  if (threadIdx.x == 0) bad = 0;
  __syncthreads();
  if (n <= 64) cholesky_registers<4>(Sm, L, Xm, n, lcol, xr, bad);
  else if (n <= 128) cholesky_registers<8>(Sm, L, Xm, n, lcol, xr, bad);
  else if (n <= 192) cholesky_registers<12>(Sm, L, Xm, n, lcol, xr, bad);
  else if (n <= 256) cholesky_registers<16>(Sm, L, Xm, n, lcol, xr, bad);
  else cholesky_global(Sm, L, Xm, n, bad);
  if (threadIdx.x == 0 && bad) status[s].status = 8;
}

// Overlap factorization for n > 256, 512 threads, X = L^{-T} as above (L in
// `work`, column-major lower; X row-major upper with its strict lower part
// zero).
//   factorization  right-looking in panels of kCholPanel columns: inside a
//                  panel column by column (the pivot row's panel entries go
//                  through LDS, every thread owns rows); after the panel the
//                  trailing lower triangle takes A -= L21 L21^T once, in
//                  16 x 16 tiles on the FP64 matrix cores (K = the panel).
//   inverse        L^{-1} by 16 x 16 blocks: the diagonal blocks by
//                  substitution (a thread per block column), then block row
//                  I from the rows above it,
//                    Linv_IJ = -Linv_II sum_{K=J}^{I-1} L_IK Linv_KJ,
//                  one block J per wavefront at a time (matrix cores, the sum
//                  staged through the wavefront's LDS tile), one barrier per
//                  block row.  Linv_ij is stored as X_ji.
// Matrix-core lane mapping (v_mfma_f64_16x16x4f64): for D = P Q^T over k,
// lane (l16, lk) supplies a = Q[l16][k], b = P[l16][k] for k = 4 q + lk and
// receives D[l16][lk + 4 r], r < 4.
constexpr int kCholPanel = 16;

__global__ __launch_bounds__(512) void k_cholesky_blocked(Topo t, const int* order, const double* S,
                                                          double* work, double* X, SysState* status) {
  const int s = order[blockIdx.x];
  const int n = t.ao_off[s + 1] - t.ao_off[s];
  const int64_t mo = t.mat_off[s];
  const double* __restrict__ Sm = S + mo;
  double* __restrict__ L = work + mo;
  double* __restrict__ Xm = X + mo;
  const int tid = threadIdx.x, nt = blockDim.x;
  const int lane = tid & 63, wv = __builtin_amdgcn_readfirstlane(tid >> 6), nw = nt >> 6;
  const int l16 = lane & 15, lk = lane >> 4;
  __shared__ double pl[kCholPanel];
  __shared__ double sh_ipiv;
  __shared__ double sS[8][16][17];
  __shared__ int bad;
  auto at = [n](int r, int c) __attribute__((always_inline)) { return (int64_t)r + (int64_t)c * n; };
  if (tid == 0) bad = 0;
  for (int64_t idx = tid; idx < (int64_t)n * n; idx += nt) L[idx] = Sm[idx];
  __syncthreads();
  for (int k0 = 0; k0 < n; k0 += kCholPanel) {
    const int jn = min(kCholPanel, n - k0);
    for (int j = k0; j < k0 + jn; ++j) {
      if (tid == 0) {
        double d = L[at(j, j)];
        if (!(d > 0.0)) {
          bad = 1;
          d = 1.0;
        }
        const double piv = sqrt(d);
        const double ipiv = 1.0 / piv;
        L[at(j, j)] = piv;
        for (int jj = j + 1; jj < k0 + jn; ++jj) pl[jj - k0] = L[at(jj, j)] * ipiv;
        sh_ipiv = ipiv;
      }
      __syncthreads();
      const double ipiv = sh_ipiv;
      for (int i = j + 1 + tid; i < n; i += nt) {
        const double v = L[at(i, j)] * ipiv;
        L[at(i, j)] = v;
        const int je = min(i, k0 + jn - 1);
        for (int jj = j + 1; jj <= je; ++jj) L[at(i, jj)] -= v * pl[jj - k0];
      }
      __syncthreads();
    }
    // trailing lower triangle, rows and columns >= c0: A -= L21 L21^T
    const int c0 = k0 + jn, mt = n - c0;
    const int nb16 = (mt + 15) >> 4;
    const int ntiles = nb16 * (nb16 + 1) / 2;
    for (int tile = wv; tile < ntiles; tile += nw) {
      int TI = (int)((sqrt(8.0 * tile + 1.0) - 1.0) * 0.5);
      while (TI * (TI + 1) / 2 > tile) --TI;
      while ((TI + 1) * (TI + 2) / 2 <= tile) ++TI;
      const int TJ = tile - TI * (TI + 1) / 2;
      const int rr = 16 * TI + l16, cc = 16 * TJ + l16;  // this lane's row of P (rows), of Q (columns)
      const int rrc = c0 + min(rr, mt - 1), ccc = c0 + min(cc, mt - 1);
      double xa[kCholPanel / 4], xb[kCholPanel / 4];
#pragma unroll
      for (int q = 0; q < kCholPanel / 4; ++q) {
        const int ci = k0 + min(4 * q + lk, jn - 1);
        xa[q] = L[at(ccc, ci)];
        xb[q] = L[at(rrc, ci)];
      }
      mfma_d4 acc = mfma_d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int q = 0; q < kCholPanel / 4; ++q) {
        const bool on = 4 * q + lk < jn;
        acc = mfma_16x16x4(on && cc < mt ? xa[q] : 0.0, on && rr < mt ? xb[q] : 0.0, acc);
      }
      // acc[r] = D[row = 16 TI + l16][col = 16 TJ + lk + 4 r]
      const int row = 16 * TI + l16;
      double old[4];
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int col = 16 * TJ + lk + 4 * r;
        old[r] = L[at(c0 + min(row, mt - 1), c0 + min(col, mt - 1))];
      }
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int col = 16 * TJ + lk + 4 * r;
        if (row < mt && col < mt && row >= col) L[at(c0 + row, c0 + col)] = old[r] - acc[r];
      }
    }
    __syncthreads();
  }
  // ---- X = L^{-T}: strict lower part of X (upper part of L^{-1}) is zero ----
  for (int64_t idx = tid; idx < (int64_t)n * n; idx += nt) {
    const int r = (int)(idx / n), c = (int)(idx - (int64_t)r * n);
    if (c < r) Xm[idx] = 0.0;
  }
  // diagonal blocks: thread (I, c) solves column c of L_II^{-1}
  const int NB = (n + 15) >> 4;
  for (int idx = tid; idx < 16 * NB; idx += nt) {
    const int I = idx >> 4, c = idx & 15, o = 16 * I, b = min(16, n - o);
    if (c >= b) continue;
    double y[16];
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      y[r] = 0.0;
      if (r < c || r >= b) continue;
      double v = r == c ? 1.0 : 0.0;
#pragma unroll
      for (int k = 0; k < r; ++k)
        if (k >= c) v -= L[at(o + r, o + k)] * y[k];
      y[r] = v / L[at(o + r, o + r)];
      Xm[(int64_t)(o + c) * n + o + r] = y[r];
    }
  }
  __syncthreads();
  // block rows I >= 1: blocks J < I, one per wavefront at a time
  for (int I = 1; I < NB; ++I) {
    const int oi = 16 * I;
    const int ri = oi + l16;  // this lane's row of block row I
    const bool rin = ri < n;
    const int ric = min(ri, n - 1);
    for (int J = wv; J < I; J += nw) {
      const int oj = 16 * J;
      mfma_d4 acc = mfma_d4{0.0, 0.0, 0.0, 0.0};
      for (int K = J; K < I; ++K) {
        const int ok = 16 * K;
        double xa[4], xb[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int k = 4 * q + lk;
          xa[q] = Xm[(int64_t)(oj + l16) * n + ok + k];  // Linv(ok + k, oj + l16)
          xb[q] = L[at(ric, ok + k)];                    // L(oi + l16, ok + k)
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) acc = mfma_16x16x4(xa[q], rin ? xb[q] : 0.0, acc);
      }
      // S[l16][lk + 4 r] -> LDS; then Linv_IJ = -Linv_II S
      double(*st)[17] = sS[wv];
#pragma unroll
      for (int r = 0; r < 4; ++r) st[l16][lk + 4 * r] = acc[r];
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      mfma_d4 acc2 = mfma_d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int k = 4 * q + lk;
        const bool kin = oi + k < n;
        const double a = st[k][l16];                                          // S[k][l16]
        const double bv = Xm[(int64_t)(oi + (kin ? k : 0)) * n + ric];       // Linv(oi + l16, oi + k)
        acc2 = mfma_16x16x4(a, rin && kin ? bv : 0.0, acc2);
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      if (rin) {
#pragma unroll
        for (int r = 0; r < 4; ++r) Xm[(int64_t)(oj + lk + 4 * r) * n + ri] = -acc2[r];
      }
    }
    __syncthreads();
  }
  if (tid == 0 && bad) status[s].status = 8;
}

// ---------------------------------------------------------------------------
// Overlap conditioning gate of the reference (eigensolver.cpp:2388-2402):
// rcond_1 = 1 / (|S|_1 est(|S^{-1}|_1)) >= 1e-12, where est is LAPACK's
// Hager-Higham 1-norm estimate (dlacn2, the iteration dpocon runs; Higham,
// ACM TOMS 14 (1988) 381-396, Algorithm 4.1): at most five power steps
// x -> sign(S^{-1} x) -> S^{-1} sign(.) -> unit vector at the largest entry,
// stopped on a repeated sign vector, no increase, or a repeated index, then
// the alternating-sign test vector x_i = (-1)^i (1 + i/(n-1)) with estimate
// 2 |S^{-1} x|_1 / (3n).  Here S^{-1} v = X (X^T v) with X = L^{-T} (row-major,
// upper triangular) from k_cholesky_inverse.  One 512-thread block per system
// whose factorization succeeded.
// ---------------------------------------------------------------------------
constexpr double kMinOverlapRcond = 1e-12;

constexpr int kRcondThreads = 512;
constexpr int kRcondWaves = kRcondThreads / 64;

struct RcondShared {
  double x[kMaxAO];
  double y[kMaxAO];
  double part[kRcondWaves][kMaxAO];  // per-wave partial column sums of X^T x
  signed char sgn[kMaxAO];
  double red[kRcondWaves];
  int ired[kRcondWaves];
};

// The same vectors for systems above kMaxAO AOs, in their global slice (x, y:
// n doubles each, sgn: n bytes); red and ired stay in LDS.
struct RcondBig {
  double* x;
  double* y;
  signed char* sgn;
  double* red;
  int* ired;
};

// block sum of |v_i|
template <typename SH>
__device__ double rc_asum(const double* v, int n, SH& sh) {
  double a = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) a += fabs(v[i]);
  a = wave_sum(a);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) sh.red[threadIdx.x >> 6] = a;
  __syncthreads();
  double t = 0.0;
  for (int w = 0; w < kRcondWaves; ++w) t += sh.red[w];
  __syncthreads();
  return t;
}

// first index of the largest |v_i| (LAPACK idamax)
template <typename SH>
__device__ int rc_iamax(const double* v, int n, SH& sh) {
  double best = -1.0;
  int bi = 0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    const double a = fabs(v[i]);
    if (a > best) { best = a; bi = i; }
  }
  for (int off = 32; off > 0; off >>= 1) {
    const double ob = __shfl_xor(best, off, 64);
    const int oi = __shfl_xor(bi, off, 64);
    if (ob > best || (ob == best && oi < bi)) { best = ob; bi = oi; }
  }
  __syncthreads();
  if ((threadIdx.x & 63) == 0) { sh.red[threadIdx.x >> 6] = best; sh.ired[threadIdx.x >> 6] = bi; }
  __syncthreads();
  best = sh.red[0];
  bi = sh.ired[0];
  for (int w = 1; w < kRcondWaves; ++w)
    if (sh.red[w] > best || (sh.red[w] == best && sh.ired[w] < bi)) { best = sh.red[w]; bi = sh.ired[w]; }
  __syncthreads();
  return bi;
}

// sh.x <- S^{-1} sh.x  (y = X^T x, then x = X y); X row-major upper triangular
// with its strict lower part stored as zero.  Both products read X by rows,
// lanes along the row, 16 / KC rows per wave at a time (their loads are in
// flight together): X^T x as per-wave partial column sums (summed in
// a fixed order), X y as one wave per row.  KC = column chunks of 64.
template <int KC>
__device__ void rc_apply_k(const double* __restrict__ Xm, int n, RcondShared& sh) {
  constexpr int U = 16 / KC;  // rows in flight per wave: 16 loads per lane
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  __syncthreads();
  double acc[KC];
#pragma unroll
  for (int m = 0; m < KC; ++m) acc[m] = 0.0;
  for (int r0 = wv; r0 < n; r0 += U * kRcondWaves) {
    double xv[U], v[U][KC];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int r = r0 + u * kRcondWaves;
      xv[u] = r < n ? sh.x[r] : 0.0;
#pragma unroll
      for (int m = 0; m < KC; ++m) {
        const int c = lane + 64 * m;
        v[u][m] = (r < n && c >= r && c < n) ? Xm[(int64_t)r * n + c] : 0.0;
      }
    }
#pragma unroll
    for (int u = 0; u < U; ++u)
#pragma unroll
      for (int m = 0; m < KC; ++m) acc[m] = fma(v[u][m], xv[u], acc[m]);
  }
#pragma unroll
  for (int m = 0; m < KC; ++m) {
    const int c = lane + 64 * m;
    if (c < n) sh.part[wv][c] = acc[m];
  }
  __syncthreads();
  for (int c = threadIdx.x; c < n; c += blockDim.x) {
    double t = 0.0;
    for (int w = 0; w < kRcondWaves; ++w) t += sh.part[w][c];
    sh.y[c] = t;
  }
  __syncthreads();
  for (int r0 = wv; r0 < n; r0 += U * kRcondWaves) {
    double a[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int r = r0 + u * kRcondWaves;
      a[u] = 0.0;
#pragma unroll
      for (int m = 0; m < KC; ++m) {
        const int c = lane + 64 * m;
        if (r < n && c >= r && c < n) a[u] = fma(Xm[(int64_t)r * n + c], sh.y[c], a[u]);
      }
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int r = r0 + u * kRcondWaves;
      const double t = wave_sum(a[u]);
      if (lane == 0 && r < n) sh.x[r] = t;
    }
  }
  __syncthreads();
}

__device__ void rc_apply(const double* __restrict__ Xm, int n, RcondShared& sh) {
  if (n <= 256) rc_apply_k<4>(Xm, n, sh);
  else if (n <= 512) rc_apply_k<8>(Xm, n, sh);
  else rc_apply_k<16>(Xm, n, sh);
}

// n > kMaxAO: y = X^T x with a thread per column (the column's entries of
// consecutive rows are adjacent across the threads), then x = X y with a
// wave per row
__device__ void rc_apply(const double* __restrict__ Xm, int n, RcondBig& sh) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  __syncthreads();
  for (int c = threadIdx.x; c < n; c += blockDim.x) {
    double a = 0.0;
    for (int r = 0; r <= c; ++r) a = fma(Xm[(int64_t)r * n + c], sh.x[r], a);
    sh.y[c] = a;
  }
  __syncthreads();
  for (int r = wv; r < n; r += kRcondWaves) {
    double a = 0.0;
    for (int c = r + lane; c < n; c += 64) a = fma(Xm[(int64_t)r * n + c], sh.y[c], a);
    a = wave_sum(a);
    if (lane == 0) sh.x[r] = a;
  }
  __syncthreads();
}

template <typename SH>
__device__ void overlap_rcond(const double* __restrict__ Sm, const double* __restrict__ Xm, int n, SH& sh,
                              SysState& status) {
  // |S|_1: S is symmetric, so the largest row sum (one wave per row, lanes
  // along the row)
  double anorm = 0.0;
  {
    const int lane = threadIdx.x & 63;
    double m = 0.0;
    for (int r = threadIdx.x >> 6; r < n; r += kRcondWaves) {
      double a = 0.0;
      for (int c = lane; c < n; c += 64) a += fabs(Sm[(int64_t)r * n + c]);
      m = fmax(m, wave_sum(a));
    }
    for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off, 64));
    if ((threadIdx.x & 63) == 0) sh.red[threadIdx.x >> 6] = m;
    __syncthreads();
    for (int w = 0; w < kRcondWaves; ++w) anorm = fmax(anorm, sh.red[w]);
    __syncthreads();
  }
  // estimate of |S^{-1}|_1
  for (int i = threadIdx.x; i < n; i += blockDim.x) sh.x[i] = 1.0 / n;
  rc_apply(Xm, n, sh);
  double est;
  if (n == 1) {
    est = fabs(sh.x[0]);
  } else {
    est = rc_asum(sh.x, n, sh);
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
      const signed char g = sh.x[i] >= 0.0 ? 1 : -1;
      sh.sgn[i] = g;
      sh.x[i] = g;
    }
    rc_apply(Xm, n, sh);
    int j = rc_iamax(sh.x, n, sh);
    for (int iter = 2;; ++iter) {
      for (int i = threadIdx.x; i < n; i += blockDim.x) sh.x[i] = i == j ? 1.0 : 0.0;
      rc_apply(Xm, n, sh);
      const double estold = est;
      est = rc_asum(sh.x, n, sh);
      __shared__ int differs;
      if (threadIdx.x == 0) differs = 0;
      __syncthreads();
      for (int i = threadIdx.x; i < n; i += blockDim.x)
        if ((sh.x[i] >= 0.0 ? 1 : -1) != sh.sgn[i]) differs = 1;
      __syncthreads();
      if (!differs || est <= estold) break;  // repeated sign vector, or no increase
      for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const signed char g = sh.x[i] >= 0.0 ? 1 : -1;
        sh.sgn[i] = g;
        sh.x[i] = g;
      }
      rc_apply(Xm, n, sh);
      const int jlast = j;
      j = rc_iamax(sh.x, n, sh);
      if (!(sh.x[jlast] != fabs(sh.x[j]) && iter < 5)) break;
    }
    // alternating-sign test vector
    for (int i = threadIdx.x; i < n; i += blockDim.x)
      sh.x[i] = (i & 1 ? -1.0 : 1.0) * (1.0 + (double)i / (double)(n - 1));
    rc_apply(Xm, n, sh);
    const double temp = 2.0 * (rc_asum(sh.x, n, sh) / (3.0 * n));
    if (temp > est) est = temp;
  }
  const double rcond = (est != 0.0 && anorm != 0.0) ? (1.0 / est) / anorm : 0.0;
  if (threadIdx.x == 0 && !(rcond >= kMinOverlapRcond && isfinite(rcond))) status.status = 8;
}

// BIG: the systems above kMaxAO AOs, with x, y and the signs in their global
// slice of `big`
template <bool BIG>
__global__ __launch_bounds__(kRcondThreads) void k_overlap_rcond(Topo t, const int* order, const double* S,
                                                       const double* X, SysState* status, double* big) {
  const int s = order[blockIdx.x];
  if (status[s].status != 0) return;
  const int n = t.ao_off[s + 1] - t.ao_off[s];
  const int64_t mo = t.mat_off[s];
  if constexpr (BIG) {
    __shared__ double red[kRcondWaves];
    __shared__ int ired[kRcondWaves];
    double* x = big + big_offset(t, s);
    RcondBig sh{x, x + n, reinterpret_cast<signed char*>(x + 2 * n), red, ired};
    overlap_rcond(S + mo, X + mo, n, sh, status[s]);
  } else {
    __shared__ RcondShared sh;
    overlap_rcond(S + mo, X + mo, n, sh, status[s]);
  }
}

// x / y via the hardware reciprocal refined by one Newton step (a few ulp,
// enough for the sign decisions of a Sturm count)
__device__ __forceinline__ double fast_div(double x, double y) {
  double r = __builtin_amdgcn_rcp(y);
  r = fma(r, fma(-y, r, 1.0), r);
  return x * r;
}

// Sturm count of T[b0:b1) - x I (the number of eigenvalues below x) together
// with L(x) = d/dx log|det(T - x I)| = sum_i 1/(x - lambda_i), from the
// three-term recurrence of the leading principal minors of T - x I
//   p_i = (d_i - x) p_{i-1} - e_{i-1}^2 p_{i-2},   p_{-1} = 1,
// and of their derivatives
//   p_i' = (d_i - x) p_{i-1}' - p_{i-1} - e_{i-1}^2 p_{i-2}',  p_{-1}' = 0.
// The count is the number of sign changes along p_{-1}, p_0, ..., p_{b-1}
// (a vanishing minor takes the sign opposite to its predecessor, as a
// vanishing pivot of T - x I = L D L^T counts as negative), and L = p'/p of
// the last minor.  The pivots of L D L^T are the ratios p_i / p_{i-1}, so this
// is the pivot count without a division per row: one FMA per row sits on the
// serial chain.  Every eight rows both minors and both derivatives are scaled
// by the same power of two (the exponent of the larger minor), which keeps
// them in range and changes neither a sign nor the ratio p'/p.
__device__ __forceinline__ int sturm_count_logd(const double* d, const double* e2, int b0, int b1,
                                                double x, double& L) {
  double pm2 = 1.0, pm1 = d[b0] - x;  // p_{i-2}, p_{i-1}
  double qm2 = 0.0, qm1 = -1.0;       // their derivatives
  int neg = pm1 <= 0.0 ? 1 : 0;       // sign of p_{i-1} (p_{-1} = 1 > 0)
  int c = neg;
  // sign bookkeeping on the bits (selects, no branches: the count must not
  // put control flow on the chain)
  auto step = [&](double di, double ei2) __attribute__((always_inline)) {
    const double dx = di - x;
    const double p = fma(dx, pm1, -ei2 * pm2);
    const double q = fma(dx, qm1, -fma(ei2, qm2, pm1));
    const unsigned long long bits = (unsigned long long)__double_as_longlong(p);
    const int sgn = (int)(bits >> 63);
    const int sn = (bits << 1) == 0ull ? (neg ^ 1) : sgn;  // a zero minor flips the sign
    c += sn ^ neg;
    neg = sn;
    pm2 = pm1;
    pm1 = p;
    qm2 = qm1;
    qm1 = q;
  };
  // load eight entries of d and e^2 ahead so that LDS latency stays off the chain
  int i = b0 + 1;
  for (; i + 7 < b1; i += 8) {
    double dv[8], ev[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      dv[u] = d[i + u];
      ev[u] = e2[i - 1 + u];
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) step(dv[u], ev[u]);
    const int ex = __builtin_amdgcn_frexp_exp(fmax(fabs(pm1), fabs(pm2)));
    pm1 = ldexp(pm1, -ex);
    pm2 = ldexp(pm2, -ex);
    qm1 = ldexp(qm1, -ex);
    qm2 = ldexp(qm2, -ex);
  }
  for (; i < b1; ++i) step(d[i], e2[i - 1]);
  L = qm1 / pm1;
  return c;
}

// Eigenvalues of T by Sturm-count bracketing with Newton steps on
// det(T - x I) once a bracket isolates exactly its eigenvalue (a step outside
// the bracket falls back to bisection).  Converged when the bracket or the
// Newton step is below 2 eps |T_b| + pivmin, the accuracy of bisection.  A
// guess (the previous SCC round's eigenvalue) first narrows the bracket to
// [guess - delta, guess + delta] where the counts allow.
//
// The lanes of the block share the eigenvalues through an LDS counter: a lane
// whose eigenvalue converged takes the next index, so a wave's cost follows
// the total work instead of its slowest eigenvalue.  One trip of the loop is
// one pass over the block.
struct EigState {
  int j, b0, b1, k, it, phase;  // phase 0/1: guess bracket ends, 2: main iteration
  int clo, chi;
  double lo, hi, x, tol, guess, dl, dh;
  double dxo;  // size of the step before the last one (Newton acceptance)
};

// Gershgorin interval and tolerance of the unreduced block [b0, b1)
__device__ __forceinline__ void block_bounds(const double* d, const double* e, int b0, int b1,
                                             double pivmin, double& lo, double& hi, double& tol) {
  double gl = 1e300, gu = -1e300;
  for (int i = b0; i < b1; ++i) {
    const double off = (i > b0 ? fabs(e[i - 1]) : 0.0) + (i + 1 < b1 ? fabs(e[i]) : 0.0);
    gl = fmin(gl, d[i] - off);
    gu = fmax(gu, d[i] + off);
  }
  const double tnorm = fmax(fabs(gl), fabs(gu));
  const double fudge = 2.0 * kEps * tnorm * (b1 - b0) + 4.0 * pivmin;
  lo = gl - fudge;
  hi = gu + fudge;
  tol = 2.0 * kEps * tnorm + pivmin;
}

// starts eigenvalue j; when T is unsplit the whole-matrix bounds are passed in
__device__ __forceinline__ bool eig_start(EigState& st, int j, const double* d, const double* e,
                                          int n, double pivmin, bool unsplit, double glo,
                                          double ghi, double gtol, bool guess_ok, double guess,
                                          double gprev, double gnext, double delta, double* lam) {
  int b0 = 0, b1 = n;
  if (!unsplit) {
    b0 = j;
    b1 = j + 1;
    while (b0 > 0 && e[b0 - 1] != 0.0) --b0;
    while (b1 < n && e[b1 - 1] != 0.0) ++b1;
  }
  if (b1 - b0 == 1) {
    lam[j] = d[j];
    return false;
  }
  st.j = j; st.b0 = b0; st.b1 = b1; st.k = j - b0; st.it = 0;
  if (unsplit) {
    st.lo = glo; st.hi = ghi; st.tol = gtol;
  } else {
    block_bounds(d, e, b0, b1, pivmin, st.lo, st.hi, st.tol);
  }
  st.clo = 0;
  st.chi = b1 - b0;
  const bool g = guess_ok && guess > st.lo && guess < st.hi;
  st.guess = guess;
  // guess bracket: at most the largest change of the last round, and at most
  // half the distance to the neighbouring guesses, so that the bracket usually
  // isolates the eigenvalue at once (Newton then starts from the guess)
  const double dmin = 4.0 * st.tol;
  st.dl = fmax(fmin(delta, 0.5 * (guess - gprev)), dmin);
  st.dh = fmax(fmin(delta, 0.5 * (gnext - guess)), dmin);
  st.phase = g ? 0 : 2;
  st.x = g ? guess - st.dl : 0.5 * (st.lo + st.hi);
  st.dxo = st.hi - st.lo;
  return true;
}

// one pass at st.x; returns true (result in `out`) when the eigenvalue has converged
__device__ __forceinline__ bool eig_step(EigState& st, const double* d, const double* e2, double& out) {
  double L;
  const int c = sturm_count_logd(d, e2, st.b0, st.b1, st.x, L);
  if (st.phase < 2) {
    // guess bracket: only narrow where the count confirms the side
    if (st.x > st.lo && st.x < st.hi) {
      if (c <= st.k) { st.lo = st.x; st.clo = c; } else { st.hi = st.x; st.chi = c; }
    }
    if (st.phase == 0) {
      st.phase = 1;
      st.x = st.guess + st.dh;
      if (st.x < st.hi) return false;
    }
    st.phase = 2;
    st.x = (st.guess > st.lo && st.guess < st.hi) ? st.guess : 0.5 * (st.lo + st.hi);
    st.dxo = st.hi - st.lo;
    return false;
  }
  if (c <= st.k) { st.lo = st.x; st.clo = c; } else { st.hi = st.x; st.chi = c; }
  if (st.hi - st.lo <= st.tol || ++st.it >= 256) {
    out = 0.5 * (st.lo + st.hi);
    return true;
  }
  // Newton only on an isolating bracket and only while it converges: the step
  // must stay inside and be less than half the step before last (otherwise
  // bisect), which rules out slow one-sided creeping
  const double dxo = st.dxo;
  if (st.clo == st.k && st.chi == st.k + 1 && L != 0.0 && isfinite(L)) {
    const double step = -1.0 / L;
    if (fabs(step) <= st.tol) {
      out = st.x + step;
      return true;
    }
    const double xn = st.x + step;
    if (xn > st.lo && xn < st.hi && fabs(step) < 0.5 * dxo) {
      st.dxo = fabs(step);
      st.x = xn;
      return false;
    }
  }
  const double mid = 0.5 * (st.lo + st.hi);
  if (mid == st.lo || mid == st.hi) {
    out = mid;
    return true;
  }
  st.dxo = fabs(mid - st.x);
  st.x = mid;
  return false;
}

// deterministic start vector in [-1, 1)
__device__ __forceinline__ double start_component(int j, int i) {
  uint64_t z = (uint64_t)(j + 1) * 0x9E3779B97F4A7C15ull + (uint64_t)(i + 1) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  return (double)(z >> 11) * (2.0 / 9007199254740992.0) - 1.0;
}

// Per-system scratch for the LU factors of one shifted block, interleaved by
// vector (element i of vector j at [i * n + j]) so that a wavefront's threads
// touch consecutive addresses.
struct EigenScratch {
  double* u0;    // pivot row diagonal
  double* u1;    // first superdiagonal
  double* u2;    // second superdiagonal (fill-in from row interchanges)
  double* mult;  // elimination multipliers
  double* piv;   // 1.0 where rows k, k+1 were interchanged
};

// Eigenvector of an isolated eigenvalue x of the unreduced block T[b0:b1) from
// its twisted factorization: forward pivots D+ (in F0) and backward pivots D-
// (in F1) of T - x I, twist index r = argmin |D+_r + D-_r - (d_r - x)|, then
// z_r = 1, z_i = -(e_i / D+_i) z_{i+1} below r and z_i = -(e_{i-1} / D-_i) z_{i-1}
// above r.  Z receives the unnormalized vector and `scale` the factor that
// normalizes it with the largest component positive (applied when the vector
// is back-transformed).  Returns false (leaving Z to be overwritten) if the
// recurrences produced a non-finite value.
__device__ __forceinline__ bool twisted_vector(const double* d, const double* e, int b0, int b1,
                                               double x, double pivmin, int n, int jj,
                                               double* __restrict__ Dp, double* __restrict__ Dm,
                                               double* __restrict__ Z, double& scale) {
  // Every loop below is a serial chain; the entries it consumes (d, e from
  // LDS, pivots from global memory) are loaded eight at a time ahead of it.
  constexpr int U = 8;
  auto at = [&](int i) __attribute__((always_inline)) { return (int64_t)i * n + jj; };
  double q = d[b0] - x;
  if (fabs(q) < pivmin) q = -pivmin;
  Dp[at(b0)] = q;
  {
    auto step = [&](int i, double di, double ei) __attribute__((always_inline)) {
      q = (di - x) - fast_div(ei * ei, q);
      if (fabs(q) < pivmin) q = -pivmin;
      Dp[at(i)] = q;
    };
    int i = b0 + 1;
    for (; i + U - 1 < b1; i += U) {
      double dv[U], ev[U];
#pragma unroll
      for (int u = 0; u < U; ++u) { dv[u] = d[i + u]; ev[u] = e[i - 1 + u]; }
#pragma unroll
      for (int u = 0; u < U; ++u) step(i + u, dv[u], ev[u]);
    }
    for (; i < b1; ++i) step(i, d[i], e[i - 1]);
  }
  q = d[b1 - 1] - x;
  if (fabs(q) < pivmin) q = -pivmin;
  Dm[at(b1 - 1)] = q;
  int r = b1 - 1;
  double gmin = fabs(Dp[at(b1 - 1)] + q - (d[b1 - 1] - x));
  {
    auto step = [&](int i, double di, double ei, double dpi) __attribute__((always_inline)) {
      q = (di - x) - fast_div(ei * ei, q);
      if (fabs(q) < pivmin) q = -pivmin;
      Dm[at(i)] = q;
      const double g = fabs(dpi + q - (di - x));
      if (g < gmin) {
        gmin = g;
        r = i;
      }
    };
    int i = b1 - 2;
    for (; i - U + 1 >= b0; i -= U) {
      double dv[U], ev[U], pv[U];
#pragma unroll
      for (int u = 0; u < U; ++u) { dv[u] = d[i - u]; ev[u] = e[i - u]; pv[u] = Dp[at(i - u)]; }
#pragma unroll
      for (int u = 0; u < U; ++u) step(i - u, dv[u], ev[u], pv[u]);
    }
    for (; i >= b0; --i) step(i, d[i], e[i], Dp[at(i)]);
  }
  double z = 1.0, nrm = 1.0, big = 1.0;
  Z[at(r)] = 1.0;
  auto put = [&](int i, double ratio) __attribute__((always_inline)) {
    z = -ratio * z;
    Z[at(i)] = z;
    nrm += z * z;
    if (fabs(z) > fabs(big)) big = z;
  };
  {
    int i = r - 1;
    for (; i - U + 1 >= b0; i -= U) {
      double rv[U];
#pragma unroll
      for (int u = 0; u < U; ++u) rv[u] = e[i - u] / Dp[at(i - u)];
#pragma unroll
      for (int u = 0; u < U; ++u) put(i - u, rv[u]);
    }
    for (; i >= b0; --i) put(i, e[i] / Dp[at(i)]);
  }
  z = 1.0;
  {
    int i = r + 1;
    for (; i + U - 1 < b1; i += U) {
      double rv[U];
#pragma unroll
      for (int u = 0; u < U; ++u) rv[u] = e[i - 1 + u] / Dm[at(i + u)];
#pragma unroll
      for (int u = 0; u < U; ++u) put(i + u, rv[u]);
    }
    for (; i < b1; ++i) put(i, e[i - 1] / Dm[at(i)]);
  }
  if (!isfinite(nrm) || !(nrm > 0.0)) return false;
  scale = (big < 0.0 ? -1.0 : 1.0) / sqrt(nrm);
  return true;
}

// ---------------------------------------------------------------------------
// Fermi occupations of one spin channel (reference eigensolver.cpp:992-1161),
// FP64 bisection on the shifted frame; one wavefront.
// ---------------------------------------------------------------------------
__device__ double fermi_sum_wave(const double* eps, int n, double ref, double mu, double kt,
                                 bool holes) {
  double acc = 0.0;
  for (int i = threadIdx.x & 63; i < n; i += 64) {
    const double a = ((eps[i] - ref) - mu) / kt;
    double f;
    if (holes) f = a >= 0.0 ? 1.0 / (exp(-a) + 1.0) : exp(a) / (1.0 + exp(a));
    else f = a >= 0.0 ? exp(-a) / (1.0 + exp(-a)) : 1.0 / (exp(a) + 1.0);
    acc += f;
  }
  return wave_sum(acc);
}

// (h, l) += x in double-double (two-sum of the high parts, then renormalized)
__device__ __forceinline__ void dd_accumulate(double& h, double& l, double x) {
  const double sm = h + x, v = sm - h;
  const double err = (h - (sm - v)) + (x - v) + l;
  const double nh = sm + err;
  l = err - (nh - sm);
  h = nh;
}

// double-double sum over the wavefront, every lane receiving lane 0's result
__device__ __forceinline__ double dd_wave_excess(double h, double l, double qt) {
  for (int off = 32; off > 0; off >>= 1) {
    const double oh = __shfl_xor(h, off, 64), ol = __shfl_xor(l, off, 64);
    dd_accumulate(h, l, oh);
    dd_accumulate(h, l, ol);
  }
  h = __shfl(h, 0, 64);
  l = __shfl(l, 0, 64);
  dd_accumulate(h, l, -qt);
  return h + l;
}

// The counts of one spin channel at mu minus the target qt, both accumulated
// in double-double over the wavefront:
//   ideal  the occupations before rounding (holes: the hole occupations
//          1 / (e^-a + 1) as such), to the accuracy of exp and the division;
//   pub    the binary64 occupations f that are published (holes: 1 - f,
//          taken exactly).
// For electrons the two take the same terms.  The reference's long-double
// main path accepts its root when both lie within 64 eps qt of the target
// (eigensolver.cpp:1135-1161); with both counts this accurate the bisection
// below applies that same acceptance at every step, so the published count
// meets the target whenever a root is accepted.
__device__ __forceinline__ void fermi_counts_wave(const double* eps, int n, double ref, double mu,
                                                  double kt, bool holes, double qt, double& ideal,
                                                  double& pub) {
  double ih = 0.0, il = 0.0, ph = 0.0, pl = 0.0;
  for (int i = threadIdx.x & 63; i < n; i += 64) {
    const double a = ((eps[i] - ref) - mu) / kt;
    const double x = exp(-fabs(a));
    const double big = 1.0 / (x + 1.0), small = x / (1.0 + x);
    const double f = a >= 0.0 ? small : big;
    if (holes) {
      dd_accumulate(ih, il, a >= 0.0 ? big : small);
      const double h = 1.0 - f;
      dd_accumulate(ph, pl, h);
      pl += (1.0 - h) - f;  // the rounding error of 1 - f, exact for 0 <= f <= 1
    } else {
      dd_accumulate(ph, pl, f);
    }
  }
  pub = dd_wave_excess(ph, pl, qt);
  ideal = holes ? dd_wave_excess(ih, il, qt) : pub;
}

// Occupations of one spin channel on one wavefront.  Returns 0 when the
// occupations in `occ` are final, 1 when the rare path (fermi_policy) must
// publish them -- no binary64 chemical potential puts both the count before
// rounding and the published count within 64 eps of the target -- and 2 when
// the count is not bracketed (a data failure), as the reference's
// compute_occupations decides (eigensolver.cpp:1026-1161).  ideal_ok: the
// root met the count.
__device__ int fermi_wave(const double* eps, int n, double nel, double kt, double* occ,
                          double& entropy, bool& ideal_ok) {
  const int lane = threadIdx.x & 63;
  entropy = 0.0;
  ideal_ok = true;
  if (kt == 0.0) {
    const int full = (int)floor(nel);
    for (int i = lane; i < n; i += 64) occ[i] = i < full ? 1.0 : (i == full ? nel - full : 0.0);
    return 0;
  }
  if (nel == 0.0) {
    for (int i = lane; i < n; i += 64) occ[i] = 0.0;
    return 0;
  }
  if (nel == (double)n) {
    for (int i = lane; i < n; i += 64) occ[i] = 1.0;
    return 0;
  }
  const bool holes = nel > 0.5 * n;
  const double qt = holes ? n - nel : nel;
  const double steps = fmax(64.0, -log(qt / n) + 8.0);
  const double ref = holes ? eps[n - 1] : eps[0];
  const double lo0 = eps[0] - ref, hi0 = eps[n - 1] - ref;
  const double margin = steps * kt + 64.0 * 2.220446049250313e-16 * fmax(1.0, fabs(hi0 - lo0));
  double lo = lo0 - margin, hi = hi0 + margin;
  {
    const double ql = fermi_sum_wave(eps, n, ref, lo, kt, holes);
    const double qh = fermi_sum_wave(eps, n, ref, hi, kt, holes);
    const bool brk = holes ? (ql >= qt && qh <= qt) : (ql <= qt && qh >= qt);
    if (!isfinite(lo) || !isfinite(hi) || !(lo < hi) || !brk) return 2;
  }
  const double tol = 64.0 * 2.220446049250313e-16 * qt;
  double xi = 0.0, xp = 0.0;
  bool met = false;
  for (int it = 0; it < 4096; ++it) {
    const double mid = lo + 0.5 * (hi - lo);
    fermi_counts_wave(eps, n, ref, mid, kt, holes, qt, xi, xp);
    if (fabs(xi) <= tol && fabs(xp) <= tol) {
      lo = hi = mid;
      met = true;
      break;
    }
    if (mid == lo || mid == hi) break;
    if ((!holes && xi < 0.0) || (holes && xi > 0.0)) lo = mid; else hi = mid;
  }
  const double mu = lo + 0.5 * (hi - lo);
  if (!met) fermi_counts_wave(eps, n, ref, mu, kt, holes, qt, xi, xp);
  ideal_ok = fabs(xi) <= tol;
  const bool pub_ok = fabs(xp) <= tol;
  double sacc = 0.0;
  for (int i = lane; i < n; i += 64) {
    const double a = ((eps[i] - ref) - mu) / kt;
    const double x = exp(-fabs(a));
    const double f = a >= 0.0 ? x / (1.0 + x) : 1.0 / (x + 1.0);
    occ[i] = f;
    if (f > 0.0 && f < 1.0) sacc -= f * log(f) + (1.0 - f) * log(1.0 - f);
  }
  entropy = wave_sum(sacc);
  return ideal_ok && pub_ok ? 0 : 1;
}

// Occupations are negligible below kOccFloor: orbitals past the last one
// above it add less than 1e-30 of a unit occupation to P and W, far below
// FP64 resolution, and get no eigenvector.
constexpr double kOccFloor = 1e-30;

// Fermi outputs of the eigen kernel, per orbital in ascending eigenvalue order.
struct FermiOut {
  double* fa;    // alpha occupations
  double* fb;    // beta occupations
  double* occ;   // fa + fb
  double* fe;    // occ * eps
  SysState* st;  // entropy and kocc
  int* kocc;     // [B] leading orbitals with occ > kOccFloor
  double kt;
};

// Eigenvalues of T (all), then the Fermi occupations, then eigenvectors for the
// kocc lowest eigenvalues only, written to Z column = ascending rank.
// From the second SCC round on, wall holds the previous round's ascending
// eigenvalues on entry; shift[s] carries the largest eigenvalue change of the
// system's previous round (written here for the next one).  BIG: the systems
// above kMaxAO AOs, with d, e, e^2, the eigenvalues and their ranks in their
// global slice of `big`.  Clusters of more than kSerialCluster close
// eigenvalues are left to k_cluster_vectors through `co`.
template <bool BIG>
__global__ __launch_bounds__(256) void k_tridiag_eigen(Topo t, const int* order,
                                                       const unsigned char* active,
                                                       const double* dall, const double* eall,
                                                       double* Zall, EigenScratch sc,
                                                       double* wall, double* shift, int round_arg,
                                                       const int* round_ptr,
                                                       double* zscale_all, FermiOut fo, double* big,
                                                       ClusterOut co) {
  const int s = order[blockIdx.x];
  if (active && !active[s]) return;
  const int n = t.ao_off[s + 1] - t.ao_off[s];
  const int64_t mo = t.mat_off[s];
  double* __restrict__ Z = Zall + mo;
  double* __restrict__ F0 = sc.u0 + mo;
  double* __restrict__ F1 = sc.u1 + mo;
  double* __restrict__ F2 = sc.u2 + mo;
  double* __restrict__ FM = sc.mult + mo;
  double* __restrict__ FP = sc.piv + mo;
  double* zscale = zscale_all + t.ao_off[s];
  __shared__ double d_l[BIG ? 1 : kMaxAO];
  __shared__ double e_l[BIG ? 1 : kMaxAO];   // split off-diagonal (negligible entries zeroed)
  __shared__ double e2_l[BIG ? 1 : kMaxAO];
  __shared__ double lam_l[BIG ? 1 : kMaxAO];
  double *d = d_l, *e = e_l, *e2 = e2_l, *lam = lam_l;
  if constexpr (BIG) {
    d = big + big_offset(t, s);  // d, e, e2, lam, rnk: 5 n <= kBigPerAO n doubles
    e = d + n;
    e2 = e + n;
    lam = e2 + n;
  }
  __shared__ double red[16];
  const int tid = threadIdx.x, nt = blockDim.x;
  if (tid == 0) co.flag[s] = 0;
  double emax2 = 0.0;
  for (int i = tid; i < n; i += nt) {
    d[i] = dall[t.ao_off[s] + i];
    double ei = i + 1 < n ? eall[t.ao_off[s] + i] : 0.0;
    e[i] = ei;
  }
  __syncthreads();
  for (int i = tid; i < n; i += nt) {
    double ei = e[i];
    if (i + 1 < n && ei * ei <= kEps * kEps * fabs(d[i] * d[i + 1]) + kSafeMin) ei = 0.0;
    e2[i] = ei * ei;
    emax2 = fmax(emax2, ei * ei);
  }
  __syncthreads();
  for (int i = tid; i < n; i += nt)
    if (e2[i] == 0.0) e[i] = 0.0;
  // block-wide max of e^2 for the pivot floor
  {
    double m2 = emax2;
    for (int off = 32; off > 0; off >>= 1) m2 = fmax(m2, __shfl_xor(m2, off, 64));
    __syncthreads();
    if ((tid & 63) == 0) red[tid >> 6] = m2;
    __syncthreads();
    m2 = 0.0;
    for (int w = 0; w < (nt >> 6); ++w) m2 = fmax(m2, red[w]);
    emax2 = m2;
    __syncthreads();
  }
  const double pivmin = kSafeMin * fmax(1.0, emax2);

  // ---- eigenvalues: lanes draw indices from a shared counter ----
  // the previous round's eigenvalues are guesses when T did not split
  __shared__ int nsplit, next_j;
  if (tid == 0) { nsplit = 0; next_j = 0; }
  __syncthreads();
  for (int i = tid; i + 1 < n; i += nt)
    if (e[i] == 0.0) atomicAdd(&nsplit, 1);
  __syncthreads();
  const bool unsplit = nsplit == 0;
  // the SCC round: the launch argument, or the device-resident counter of
  // rounds enqueued ahead of the host (graph-captured rounds)
  const int round = round_ptr ? *round_ptr : round_arg;
  const bool have_old = round >= 1 && unsplit;
  const bool guess_ok = round >= 2 && unsplit;
  const double delta = guess_ok ? 2.0 * shift[s] : 0.0;
  // whole-matrix Gershgorin bounds (used when T is unsplit)
  __shared__ double gbound[3];
  if (unsplit) {
    double gl = 1e300, gu = -1e300;
    for (int i = tid; i < n; i += nt) {
      const double off = (i > 0 ? fabs(e[i - 1]) : 0.0) + (i + 1 < n ? fabs(e[i]) : 0.0);
      gl = fmin(gl, d[i] - off);
      gu = fmax(gu, d[i] + off);
    }
    for (int off = 32; off > 0; off >>= 1) {
      gl = fmin(gl, __shfl_xor(gl, off, 64));
      gu = fmax(gu, __shfl_xor(gu, off, 64));
    }
    if ((tid & 63) == 0) { red[tid >> 6] = gl; red[8 + (tid >> 6)] = gu; }
    __syncthreads();
    if (tid == 0) {
      for (int w = 1; w < (nt >> 6); ++w) { gl = fmin(gl, red[w]); gu = fmax(gu, red[8 + w]); }
      const double tnorm = fmax(fabs(gl), fabs(gu));
      const double fudge = 2.0 * kEps * tnorm * n + 4.0 * pivmin;
      gbound[0] = gl - fudge;
      gbound[1] = gu + fudge;
      gbound[2] = 2.0 * kEps * tnorm + pivmin;
    }
    __syncthreads();
  }
  double dmax = 0.0;
  {
    EigState st;
    bool busy = false;
    for (;;) {
      if (!busy) {
        const int j = atomicAdd(&next_j, 1);
        if (j >= n) break;
        const double* wo = wall + t.ao_off[s];
        const double g = have_old ? wo[j] : 0.0;
        const double gp = (have_old && j > 0) ? wo[j - 1] : -1e300;
        const double gn = (have_old && j + 1 < n) ? wo[j + 1] : 1e300;
        busy = eig_start(st, j, d, e, n, pivmin, unsplit, gbound[0], gbound[1], gbound[2], guess_ok,
                         g, gp, gn, delta, lam);
        if (!busy) continue;
      }
      double x;
      if (eig_step(st, d, e2, x)) {
        lam[st.j] = x;
        if (have_old) dmax = fmax(dmax, fabs(x - wall[t.ao_off[s] + st.j]));
        busy = false;
      }
    }
  }
  // largest eigenvalue change of this round (block max), for the next round's brackets
  for (int off = 32; off > 0; off >>= 1) dmax = fmax(dmax, __shfl_xor(dmax, off, 64));
  __syncthreads();
  if ((tid & 63) == 0) red[tid >> 6] = dmax;
  __syncthreads();
  if (tid == 0) {
    double m = 0.0;
    for (int w = 0; w < (nt >> 6); ++w) m = fmax(m, red[w]);
    // first round: no change measured yet, take a generous bracket
    shift[s] = have_old ? m : 1.0;
  }
  __syncthreads();

  // ---- ascending rank (stable on ties) and the sorted spectrum ----
  __shared__ int rnk_l[BIG ? 1 : kMaxAO];
  int* rnk = BIG ? reinterpret_cast<int*>(lam + n) : rnk_l;
  __shared__ int kocc_sh;
  double* srt = e2;  // e2 is not needed after the eigenvalues
  for (int j = tid; j < n; j += nt) {
    const double lj = lam[j];
    int r = 0;
    for (int i = 0; i < n; ++i) {
      const double li = lam[i];
      r += (li < lj) || (li == lj && i < j);
    }
    rnk[j] = r;
    srt[r] = lj;
    wall[t.ao_off[s] + r] = lj;
  }
  __syncthreads();
  // ---- occupations: alpha on the first wavefront, beta on the second (a
  // closed shell solves once and copies) ----
  __shared__ int chan_r[2];
  __shared__ double chan_s[2];
  const bool same_channels = t.nalpha[s] == t.nbeta[s];
  if (tid < 128 && (tid < 64 || !same_channels)) {
    const int w = tid >> 6;
    const int o = t.ao_off[s];
    const double nel = w == 0 ? t.nalpha[s] : t.nbeta[s];
    double* f = (w == 0 ? fo.fa : fo.fb) + o;
    double sw;
    bool ideal;
    int r = fermi_wave(srt, n, nel, fo.kt, f, sw, ideal);
    // the reference's binary64 rare path, serial on lane 0
    if (r == 1) {
      if ((tid & 63) == 0) r = fermi_policy(srt, n, nel, fo.kt, f, sw) ? 0 : 2;
      __threadfence_block();
      r = __shfl(r, 0, 64);
      sw = __shfl(sw, 0, 64);
    }
    if ((tid & 63) == 0) {
      chan_r[w] = r;
      chan_s[w] = sw;
    }
  }
  __syncthreads();
  if (tid < 64) {
    const int o = t.ao_off[s];
    const int rb = same_channels ? chan_r[0] : chan_r[1];
    const double sb = same_channels ? chan_s[0] : chan_s[1];
    if ((chan_r[0] == 2 || rb == 2) && tid == 0) fo.st[s].status = 8;
    int last = 0;
    for (int i = tid; i < n; i += 64) {
      const double fa = fo.fa[o + i];
      const double fb = same_channels ? fa : fo.fb[o + i];
      if (same_channels) fo.fb[o + i] = fa;
      const double f = fa + fb;
      fo.occ[o + i] = f;
      fo.fe[o + i] = f * srt[i];
      if (f > kOccFloor) last = max(last, i + 1);
    }
    for (int off = 32; off > 0; off >>= 1) last = max(last, __shfl_xor(last, off, 64));
    if (tid == 0) {
      fo.st[s].entropy = chan_s[0] + sb;
      fo.st[s].kocc = last;
      fo.kocc[s] = last;
      kocc_sh = last;
    }
  }
  __syncthreads();
  const int kocc = kocc_sh;

  // ---- eigenvectors: thread per cluster start ----
  // A cluster is needed when its first (lowest) member is below kocc; every
  // member of a needed cluster is computed so the cluster stays orthogonal.
  // an unsplit T is one block [0, n) whose 1-norm is the block-wide maximum
  // row sum; split blocks are found and measured by the thread that needs them
  __shared__ double tnorm1;
  if (unsplit) {
    double m = 0.0;
    for (int i = tid; i < n; i += nt)
      m = fmax(m, fabs(d[i]) + (i > 0 ? fabs(e[i - 1]) : 0.0) + (i + 1 < n ? fabs(e[i]) : 0.0));
    for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off, 64));
    if ((tid & 63) == 0) red[tid >> 6] = m;
    __syncthreads();
    if (tid == 0) {
      for (int w = 1; w < (nt >> 6); ++w) m = fmax(m, red[w]);
      tnorm1 = m;
    }
    __syncthreads();
  }
  for (int j = tid; j < n; j += nt) {
    if (rnk[j] >= kocc) continue;
    int b0 = j, b1 = j + 1;
    double onenrm = 0.0;
    if (unsplit) {
      b0 = 0;
      b1 = n;
      onenrm = tnorm1;
    } else {
      while (b0 > 0 && e[b0 - 1] != 0.0) --b0;
      while (b1 < n && e[b1 - 1] != 0.0) ++b1;
      for (int i = b0; i < b1; ++i)
        onenrm = fmax(onenrm, fabs(d[i]) + (i > b0 ? fabs(e[i - 1]) : 0.0) + (i + 1 < b1 ? fabs(e[i]) : 0.0));
    }
    const int bn = b1 - b0;
    const double gtol = kClusterTol * fmax(onenrm, kSafeMin);
    if (j > b0 && lam[j] - lam[j - 1] <= gtol) continue;  // not a cluster start
    int g1 = j + 1;
    while (g1 < b1 && lam[g1] - lam[g1 - 1] <= gtol) ++g1;
    if (g1 - j > kSerialCluster) {  // the whole block orthogonalizes it (k_cluster_vectors)
      const int o = t.ao_off[s];
      for (int jj = j; jj < g1; ++jj) {
        co.lam[o + jj] = lam[jj];
        co.rank[o + jj] = rnk[jj];
      }
      co.end[o + j] = g1;
      co.flag[s] = 1;
      continue;
    }
    const double ptol = kEps * fmax(onenrm, kSafeMin);       // smallest admissible pivot
    const double pertol = 10.0 * kEps * fmax(onenrm, kSafeMin);
    double xprev = 0.0;
    for (int jj = j; jj < g1; ++jj) {
      const int cj = rnk[jj];  // Z column of eigenvalue jj
      // rows outside the block are zero
      for (int i = 0; i < b0; ++i) Z[(int64_t)i * n + cj] = 0.0;
      for (int i = b1; i < n; ++i) Z[(int64_t)i * n + cj] = 0.0;
      zscale[cj] = 1.0;
      if (bn == 1) {
        Z[(int64_t)b0 * n + cj] = 1.0;
        continue;
      }
      double x = lam[jj];
      if (jj > j && x - xprev < pertol) x = xprev + pertol;
      xprev = x;
      if (g1 - j == 1) {
        double f;
        if (twisted_vector(d, e, b0, b1, x, pivmin, n, cj, F0, F1, Z, f)) {
          zscale[cj] = f;
          continue;
        }
      }
      // LU with partial pivoting of T_b - x I
      double c0 = d[b0] - x, c1 = e[b0], c2 = 0.0;
      for (int k = 0; k < bn - 1; ++k) {
        const int gi = b0 + k;
        const double n0 = e[gi], n1 = d[gi + 1] - x, n2 = (k + 2 < bn) ? e[gi + 1] : 0.0;
        const int64_t at = (int64_t)gi * n + cj;
        if (fabs(c0) >= fabs(n0)) {
          const double m = c0 != 0.0 ? n0 / c0 : 0.0;
          F0[at] = c0; F1[at] = c1; F2[at] = c2; FM[at] = m; FP[at] = 0.0;
          c0 = n1 - m * c1;
          c1 = n2 - m * c2;
        } else {
          const double m = c0 / n0;
          F0[at] = n0; F1[at] = n1; F2[at] = n2; FM[at] = m; FP[at] = 1.0;
          const double t0 = c1 - m * n1, t1 = c2 - m * n2;
          c0 = t0;
          c1 = t1;
        }
        c2 = 0.0;
      }
      {
        const int64_t at = (int64_t)(b1 - 1) * n + cj;
        F0[at] = c0; F1[at] = 0.0; F2[at] = 0.0; FM[at] = 0.0; FP[at] = 0.0;
      }
      for (int i = b0; i < b1; ++i) Z[(int64_t)i * n + cj] = start_component(jj, i - b0);
      for (int it = 0; it <= kInverseIterations; ++it) {
        // orthogonalize against the cluster members already finished
        for (int p = j; p < jj; ++p) {
          double dot = 0.0;
          for (int i = b0; i < b1; ++i) dot += Z[(int64_t)i * n + rnk[p]] * Z[(int64_t)i * n + cj];
          for (int i = b0; i < b1; ++i) Z[(int64_t)i * n + cj] -= dot * Z[(int64_t)i * n + rnk[p]];
        }
        if (it == kInverseIterations) break;
        // scale to unit max-norm, then solve (T_b - x I) y = r
        double rmax = 0.0;
        for (int i = b0; i < b1; ++i) rmax = fmax(rmax, fabs(Z[(int64_t)i * n + cj]));
        const double rs = rmax > 0.0 ? 1.0 / rmax : 1.0;
        for (int i = b0; i < b1; ++i) Z[(int64_t)i * n + cj] *= rs;
        // forward elimination; the running entry stays in a register (it is
        // the one stored one row earlier), so only independent loads remain
        {
          double cur = Z[(int64_t)b0 * n + cj];
          for (int k = 0; k < bn - 1; ++k) {
            const int64_t at = (int64_t)(b0 + k) * n + cj;
            double rk = cur, rn = Z[at + n];
            if (FP[at] != 0.0) { const double tmp = rk; rk = rn; rn = tmp; }
            Z[at] = rk;
            cur = rn - FM[at] * rk;
          }
          Z[(int64_t)(b1 - 1) * n + cj] = cur;
        }
        // back substitution with the two entries below in registers
        {
          double z1 = 0.0, z2 = 0.0;  // solution entries k + 1, k + 2
          for (int k = bn - 1; k >= 0; --k) {
            const int64_t at = (int64_t)(b0 + k) * n + cj;
            double r = Z[at];
            if (k + 1 < bn) r -= F1[at] * z1;
            if (k + 2 < bn) r -= F2[at] * z2;
            double u = F0[at];
            if (fabs(u) < ptol) u = u < 0.0 ? -ptol : ptol;
            const double zk = r / u;
            Z[at] = zk;
            z2 = z1;
            z1 = zk;
          }
        }
        // normalize
        double nrm = 0.0, ymax = 0.0;
        for (int i = b0; i < b1; ++i) ymax = fmax(ymax, fabs(Z[(int64_t)i * n + cj]));
        const double ys = ymax > 0.0 ? 1.0 / ymax : 1.0;
        for (int i = b0; i < b1; ++i) {
          const double y = Z[(int64_t)i * n + cj] * ys;
          nrm += y * y;
        }
        const double sc2 = ys / sqrt(nrm);
        for (int i = b0; i < b1; ++i) Z[(int64_t)i * n + cj] *= sc2;
      }
      // final normalization; sign: largest-magnitude component positive
      double nrm = 0.0, big = 0.0;
      for (int i = b0; i < b1; ++i) {
        const double y = Z[(int64_t)i * n + cj];
        nrm += y * y;
        if (fabs(y) > fabs(big)) big = y;
      }
      const double f = (big < 0.0 ? -1.0 : 1.0) / sqrt(nrm);
      for (int i = b0; i < b1; ++i) Z[(int64_t)i * n + cj] *= f;
    }
  }
}

// Block-wide maximum (blockDim.x <= 1024; red holds blockDim.x / 64 doubles).
__device__ __forceinline__ double block_max(double v, double* red) {
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
  __syncthreads();
  if (lane == 0) red[wid] = v;
  __syncthreads();
  double m = red[0];
  for (int i = 1; i < nw; ++i) m = fmax(m, red[i]);
  __syncthreads();
  return m;
}

// Eigenvectors of the clusters k_tridiag_eigen hands over (more than
// kSerialCluster members), one 256-thread block per system with such a
// cluster.  The vectors follow the serial path's inverse iteration: the same
// shift perturbation, partially pivoted LU of T_b - x I and start vector, the
// same iteration count, normalization and sign.  The orthogonalization
// against the cluster's earlier vectors runs on the whole block as classical
// Gram-Schmidt applied twice, over chunks of kCoopChunk earlier vectors (their
// dots a thread per vector, the update a thread per row); only the LU and the
// two substitutions (O(n) per vector) stay on one thread.  T's splitting, the
// block norms are recomputed from d and e as k_tridiag_eigen forms them.
constexpr int kCoopChunk = 256;

__global__ __launch_bounds__(256) void k_cluster_vectors(Topo t, const int* order, const unsigned char* active,
                                                         const double* dall, const double* eall, double* Zall,
                                                         EigenScratch sc, double* zscale_all, ClusterOut co) {
  const int s = order[blockIdx.x];
  if ((active && !active[s]) || !co.flag[s]) return;
  const int o = t.ao_off[s], n = t.ao_off[s + 1] - o;
  const int64_t mo = t.mat_off[s];
  double* __restrict__ Z = Zall + mo;
  double* __restrict__ F0 = sc.u0 + mo;
  double* __restrict__ F1 = sc.u1 + mo;
  double* __restrict__ F2 = sc.u2 + mo;
  double* __restrict__ FM = sc.mult + mo;
  double* __restrict__ FP = sc.piv + mo;
  double* zscale = zscale_all + o;
  const double* d = dall + o;
  const int tid = threadIdx.x, nt = blockDim.x;
  __shared__ double red[16];
  __shared__ int ired[16];
  __shared__ double dots[kCoopChunk];
  // off-diagonal after the split (negligible entries zeroed)
  auto esp = [&](int i) -> double {
    if (i + 1 >= n) return 0.0;
    double ei = eall[o + i];
    if (ei * ei <= kEps * kEps * fabs(d[i] * d[i + 1]) + kSafeMin) ei = 0.0;
    return ei * ei == 0.0 ? 0.0 : ei;
  };
  double rows1 = 0.0, splits = 0.0;
  for (int i = tid; i < n; i += nt) {
    const double ei = esp(i), ep = i > 0 ? esp(i - 1) : 0.0;
    rows1 = fmax(rows1, fabs(d[i]) + fabs(ep) + fabs(ei));
    if (i + 1 < n && ei == 0.0) splits += 1.0;
  }
  const double tnorm1 = block_max(rows1, red);
  const bool unsplit = block_sum(splits, red) == 0.0;
  for (int j = 0; j < n; ++j) {
    const int g1 = co.end[o + j];
    if (g1 == 0) continue;
    int b0 = j, b1 = j + 1;
    double onenrm = 0.0;
    if (unsplit) {
      b0 = 0;
      b1 = n;
      onenrm = tnorm1;
    } else {
      while (b0 > 0 && esp(b0 - 1) != 0.0) --b0;
      while (b1 < n && esp(b1 - 1) != 0.0) ++b1;
      for (int i = b0; i < b1; ++i)
        onenrm = fmax(onenrm, fabs(d[i]) + (i > b0 ? fabs(esp(i - 1)) : 0.0) + (i + 1 < b1 ? fabs(esp(i)) : 0.0));
    }
    const int bn = b1 - b0;
    const double ptol = kEps * fmax(onenrm, kSafeMin);
    const double pertol = 10.0 * kEps * fmax(onenrm, kSafeMin);
    double xprev = 0.0;
    for (int jj = j; jj < g1; ++jj) {
      const int cj = co.rank[o + jj];
      for (int i = tid; i < n; i += nt)
        if (i < b0 || i >= b1) Z[(int64_t)i * n + cj] = 0.0;
      if (tid == 0) zscale[cj] = 1.0;
      double x = co.lam[o + jj];
      if (jj > j && x - xprev < pertol) x = xprev + pertol;
      xprev = x;
      if (bn == 1) {
        if (tid == 0) Z[(int64_t)b0 * n + cj] = 1.0;
        __syncthreads();
        continue;
      }
      if (tid == 0) {  // LU with partial pivoting of T_b - x I
        double c0 = d[b0] - x, c1 = esp(b0), c2 = 0.0;
        for (int k = 0; k < bn - 1; ++k) {
          const int gi = b0 + k;
          const double n0 = esp(gi), n1 = d[gi + 1] - x, n2 = (k + 2 < bn) ? esp(gi + 1) : 0.0;
          const int64_t at = (int64_t)gi * n + cj;
          if (fabs(c0) >= fabs(n0)) {
            const double m = c0 != 0.0 ? n0 / c0 : 0.0;
            F0[at] = c0; F1[at] = c1; F2[at] = c2; FM[at] = m; FP[at] = 0.0;
            c0 = n1 - m * c1;
            c1 = n2 - m * c2;
          } else {
            const double m = c0 / n0;
            F0[at] = n0; F1[at] = n1; F2[at] = n2; FM[at] = m; FP[at] = 1.0;
            const double t0 = c1 - m * n1, t1 = c2 - m * n2;
            c0 = t0;
            c1 = t1;
          }
          c2 = 0.0;
        }
        const int64_t at = (int64_t)(b1 - 1) * n + cj;
        F0[at] = c0; F1[at] = 0.0; F2[at] = 0.0; FM[at] = 0.0; FP[at] = 0.0;
      }
      for (int i = b0 + tid; i < b1; i += nt) Z[(int64_t)i * n + cj] = start_component(jj, i - b0);
      __syncthreads();
      for (int it = 0; it <= kInverseIterations; ++it) {
        // orthogonalize against the cluster members already finished (CGS twice)
        for (int pass = 0; pass < 2; ++pass)
          for (int p0 = j; p0 < jj; p0 += kCoopChunk) {
            const int pe = min(jj, p0 + kCoopChunk);
            for (int p = p0 + tid; p < pe; p += nt) {
              const int cp = co.rank[o + p];
              double dot = 0.0;
              for (int i = b0; i < b1; ++i) dot = fma(Z[(int64_t)i * n + cp], Z[(int64_t)i * n + cj], dot);
              dots[p - p0] = dot;
            }
            __syncthreads();
            for (int i = b0 + tid; i < b1; i += nt) {
              double z = Z[(int64_t)i * n + cj];
              for (int p = p0; p < pe; ++p) z = fma(-dots[p - p0], Z[(int64_t)i * n + co.rank[o + p]], z);
              Z[(int64_t)i * n + cj] = z;
            }
            __syncthreads();
          }
        if (it == kInverseIterations) break;
        // scale to unit max-norm, then solve (T_b - x I) y = r
        double rmax = 0.0;
        for (int i = b0 + tid; i < b1; i += nt) rmax = fmax(rmax, fabs(Z[(int64_t)i * n + cj]));
        rmax = block_max(rmax, red);
        const double rs = rmax > 0.0 ? 1.0 / rmax : 1.0;
        for (int i = b0 + tid; i < b1; i += nt) Z[(int64_t)i * n + cj] *= rs;
        __syncthreads();
        if (tid == 0) {
          double cur = Z[(int64_t)b0 * n + cj];
          for (int k = 0; k < bn - 1; ++k) {
            const int64_t at = (int64_t)(b0 + k) * n + cj;
            double rk = cur, rn = Z[at + n];
            if (FP[at] != 0.0) { const double tmp = rk; rk = rn; rn = tmp; }
            Z[at] = rk;
            cur = rn - FM[at] * rk;
          }
          Z[(int64_t)(b1 - 1) * n + cj] = cur;
          double z1 = 0.0, z2 = 0.0;
          for (int k = bn - 1; k >= 0; --k) {
            const int64_t at = (int64_t)(b0 + k) * n + cj;
            double r = Z[at];
            if (k + 1 < bn) r -= F1[at] * z1;
            if (k + 2 < bn) r -= F2[at] * z2;
            double u = F0[at];
            if (fabs(u) < ptol) u = u < 0.0 ? -ptol : ptol;
            const double zk = r / u;
            Z[at] = zk;
            z2 = z1;
            z1 = zk;
          }
        }
        __syncthreads();
        // normalize
        double ymax = 0.0;
        for (int i = b0 + tid; i < b1; i += nt) ymax = fmax(ymax, fabs(Z[(int64_t)i * n + cj]));
        ymax = block_max(ymax, red);
        const double ys = ymax > 0.0 ? 1.0 / ymax : 1.0;
        double nrm = 0.0;
        for (int i = b0 + tid; i < b1; i += nt) {
          const double y = Z[(int64_t)i * n + cj] * ys;
          nrm += y * y;
        }
        nrm = block_sum(nrm, red);
        const double sc2 = ys / sqrt(nrm);
        for (int i = b0 + tid; i < b1; i += nt) Z[(int64_t)i * n + cj] *= sc2;
        __syncthreads();
      }
      // final normalization; sign: the first largest-magnitude component positive
      double nrm = 0.0, big = 0.0;
      int bi = n;
      for (int i = b0 + tid; i < b1; i += nt) {
        const double y = Z[(int64_t)i * n + cj];
        nrm += y * y;
        if (fabs(y) > fabs(big)) { big = y; bi = i; }
      }
      nrm = block_sum(nrm, red);
      for (int off = 32; off > 0; off >>= 1) {
        const double ob = __shfl_xor(big, off, 64);
        const int oi = __shfl_xor(bi, off, 64);
        if (fabs(ob) > fabs(big) || (fabs(ob) == fabs(big) && oi < bi)) { big = ob; bi = oi; }
      }
      if ((tid & 63) == 0) { red[tid >> 6] = big; ired[tid >> 6] = bi; }
      __syncthreads();
      big = red[0];
      bi = ired[0];
      for (int w = 1; w < (nt >> 6); ++w)
        if (fabs(red[w]) > fabs(big) || (fabs(red[w]) == fabs(big) && ired[w] < bi)) { big = red[w]; bi = ired[w]; }
      __syncthreads();
      const double f = (big < 0.0 ? -1.0 : 1.0) / sqrt(nrm);
      for (int i = b0 + tid; i < b1; i += nt) Z[(int64_t)i * n + cj] *= f;
      __syncthreads();
    }
    if (tid == 0) co.end[o + j] = 0;  // consumed
  }
}

// One column block of one system: Y[:, j0 .. j0+cw) = Q Z[:, j0 .. j0+cw), for
// the columns below the system's kocc (Z holds eigenvectors by ascending rank).
struct BackTask {
  int sys;
  int j0;
  int cw;
  int pad;
};

// 512 threads.  n <= 256: the 64-column block lives in registers (thread
// (c, rg) = (tid % 64, tid / 64) holds rows rg, rg + 8, ...) and the
// reflectors are applied kBackChunk at a time in compact WY form,
//   H_{k_7} ... H_{k_0} = I - V T V^T,   V = [v_{k_0} .. v_{k_7}],
// with k_u = n-3 - ch kBackChunk - u and T lower triangular (wy_factors):
// one LDS reduction (V^T z for the whole chunk) and two barriers per chunk.
// Reflectors are staged in LDS zero-padded to absolute row indices, so every
// thread applies every reflector without masks; the next chunk is fetched into
// registers while the current one is applied.  n > 256: the column block lives
// in dynamic LDS ((n cw + 2 n + 512) doubles) and reflectors are applied one at
// a time.
constexpr int kBackThreads = 512;
constexpr int kBackRegRows = 32;  // rows per thread of the register path (n <= 256)
constexpr size_t kBackStaticLds =
    (2 * kBackChunk * 256 + kBackChunk * kBackThreads + kBackChunk * 64) * sizeof(double);

__device__ __forceinline__ void backtransform_registers(const double* __restrict__ A,
                                                        const double* __restrict__ Z,
                                                        const double* __restrict__ Tf,
                                                        double* __restrict__ Y, int kocc,
                                                        const double* zscale, int n, int j0,
                                                        double* ring, double* part, double* wf) {
  // thread (w, cg, rl): wavefront w, column group cg = lane / 4 (columns
  // j0 + 4 cg + j, j < 4), row lane rl = lane % 4; rows rgp + 32 q (q < 8) with
  // rgp = 4 w + rl.  Each reflector entry read from LDS feeds four columns.
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int cg = lane >> 2, rl = lane & 3, rgp = 4 * w + rl;
  constexpr int kRows = 8, kCols = 4;
  double z[kRows][kCols];
#pragma unroll
  for (int j = 0; j < kCols; ++j) {
    const int c = j0 + kCols * cg + j;
    const bool ok = c < kocc;
    const double f = ok ? zscale[c] : 0.0;
#pragma unroll
    for (int q = 0; q < kRows; ++q) {
      const int i = rgp + 32 * q;
      z[q][j] = (ok && i < n) ? f * Z[(int64_t)i * n + c] : 0.0;
    }
  }
  // chunk ch holds k = n-3 - ch*kBackChunk - u at ring slot
  // ((ch & 1) * kBackChunk + u) * 256, zero-padded to absolute rows; its T
  // factor at Tf[64 ch].
  constexpr int kVals = kBackChunk * 256 / kBackThreads;
  constexpr int kTT = kBackChunk * kBackChunk;
  auto fetch = [&](int ch, double (&r)[kVals]) __attribute__((always_inline)) {
#pragma unroll
    for (int v = 0; v < kVals; ++v) {
      const int x = tid + kBackThreads * v, u = x >> 8, i = x & 255;
      const int k = n - 3 - ch * kBackChunk - u;
      r[v] = (k >= 0 && i > k && i < n) ? A[i + (int64_t)k * n] : 0.0;
    }
  };
  auto stage = [&](int ch, const double (&r)[kVals]) __attribute__((always_inline)) {
#pragma unroll
    for (int v = 0; v < kVals; ++v) ring[(ch & 1) * kBackChunk * 256 + tid + kBackThreads * v] = r[v];
  };
  const int nref = n - 2;  // reflectors k = 0 .. n-3
  double rn[kVals];
  if (nref > 0) {
    fetch(0, rn);
    stage(0, rn);
  }
  __syncthreads();
  for (int ch = 0; ch * kBackChunk < nref; ++ch) {
    const bool more = (ch + 1) * kBackChunk < nref;
    if (more) fetch(ch + 1, rn);  // in flight while this chunk is applied
    const double* hv = ring + (ch & 1) * kBackChunk * 256;
    // T is read-only and wave-uniform: the constant address space puts it in
    // SGPRs through scalar loads
    const __attribute__((address_space(4))) double* tt =
        (const __attribute__((address_space(4))) double*)(Tf + ch * kTT);
    // V^T z: per reflector, this thread's rows, summed over the four row lanes
    // of the column group (DPP quad permutations), then over the wavefronts
#pragma unroll
    for (int u = 0; u < kBackChunk; ++u) {
      double p[kCols] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int q = 0; q < kRows; ++q) {
        const double h = hv[u * 256 + rgp + 32 * q];
#pragma unroll
        for (int j = 0; j < kCols; ++j) p[j] = fma(h, z[q][j], p[j]);
      }
#pragma unroll
      for (int j = 0; j < kCols; ++j) {
        p[j] += dpp_move<0xb1, 0xf>(p[j]);  // lanes 0<->1, 2<->3
        p[j] += dpp_move<0x4e, 0xf>(p[j]);  // lanes 0<->2, 1<->3
      }
      if (rl == 0) {
#pragma unroll
        for (int j = 0; j < kCols; ++j) part[(u * 8 + w) * 64 + kCols * cg + j] = p[j];
      }
    }
    __syncthreads();
    // w = V^T z for reflector u, column c (thread (u, c))
    const int ur = tid >> 6, cr = tid & 63;
    {
      double acc = 0.0;
#pragma unroll
      for (int g = 0; g < 8; ++g) acc += part[(ur * 8 + g) * 64 + cr];
      wf[ur * 64 + cr] = acc;
    }
    __syncthreads();
    // y = T w (T lower triangular), thread (u, c); y goes to part[u][0][c]
    {
      double acc = 0.0;
      for (int i = 0; i <= ur; ++i) acc = fma(tt[ur * kBackChunk + i], wf[i * 64 + cr], acc);
      part[ur * 8 * 64 + cr] = acc;
    }
    __syncthreads();
    // z -= V y, one reflector at a time (only its four y values live)
#pragma unroll
    for (int u = 0; u < kBackChunk; ++u) {
      double yj[kCols];
#pragma unroll
      for (int j = 0; j < kCols; ++j) yj[j] = part[u * 8 * 64 + kCols * cg + j];
#pragma unroll
      for (int q = 0; q < kRows; ++q) {
        const double h = hv[u * 256 + rgp + 32 * q];
#pragma unroll
        for (int j = 0; j < kCols; ++j) z[q][j] = fma(-h, yj[j], z[q][j]);
      }
      // Empty asm with a memory clobber pins the program order, so that only
      // this reflector's entries are in flight (hoisting all of them spills).
#pragma unroll
      for (int q = 0; q < kRows; ++q)
        asm volatile("" : "+v"(z[q][0]), "+v"(z[q][1]), "+v"(z[q][2]), "+v"(z[q][3])::"memory");
    }
    if (more) stage(ch + 1, rn);
    __syncthreads();
  }
#pragma unroll
  for (int j = 0; j < kCols; ++j) {
    const int c = j0 + kCols * cg + j;
    if (c >= kocc) continue;
#pragma unroll
    for (int q = 0; q < kRows; ++q) {
      const int i = rgp + 32 * q;
      if (i < n) Y[(int64_t)i * n + c] = z[q][j];
    }
  }
}

__device__ __forceinline__ void backtransform_lds(const double* __restrict__ A,
                                                  const double* __restrict__ Z,
                                                  double* __restrict__ Y, int kocc,
                                                  const double* zscale, int n, int j0, int cw,
                                                  double* lds) {
  double* Zb = lds;                 // [n][cw]
  double* hvb = lds + n * cw;       // [2][n]
  double* part = hvb + 2 * n;       // [512]
  const int tid = threadIdx.x, nt = blockDim.x;
  for (int idx = tid; idx < n * cw; idx += nt) {
    const int i = idx / cw, c = idx % cw;
    Zb[idx] = (j0 + c < kocc) ? zscale[j0 + c] * Z[(int64_t)i * n + j0 + c] : 0.0;
  }
  const int c = tid % cw, rg = tid / cw, nrg = nt / cw;
  // reflector k occupies A[k+1 .. n-1, k] (column-major); m = n - k - 1 <= 1023
  int k = n - 3;
  if (k >= 0) {
    double* hv = hvb + (k & 1) * n;
    for (int i = tid; i < n - k - 1; i += nt) hv[i] = A[(k + 1 + i) + (int64_t)k * n];
  }
  __syncthreads();
  for (; k >= 0; --k) {
    const int m = n - k - 1;
    const double* hv = hvb + (k & 1) * n;
    double r0 = 0.0, r1 = 0.0;
    const bool more = k >= 1;
    if (more) {
      const int kn = k - 1;
      if (tid < m + 1) r0 = A[(kn + 1 + tid) + (int64_t)kn * n];
      if (tid + nt < m + 1) r1 = A[(kn + 1 + tid + nt) + (int64_t)kn * n];
    }
    double p = 0.0;
    for (int i = rg; i < m; i += nrg) p += hv[i] * Zb[(k + 1 + i) * cw + c];
    part[rg * cw + c] = p;
    __syncthreads();
    double dot = 0.0;
    for (int w = 0; w < nrg; ++w) dot += part[w * cw + c];
    dot *= 2.0;
    for (int i = rg; i < m; i += nrg) Zb[(k + 1 + i) * cw + c] -= dot * hv[i];
    if (more) {
      double* hn = hvb + ((k - 1) & 1) * n;
      if (tid < m + 1) hn[tid] = r0;
      if (tid + nt < m + 1) hn[tid + nt] = r1;
    }
    __syncthreads();
  }
  for (int idx = tid; idx < n * cw; idx += nt) {
    const int i = idx / cw, cc = idx % cw;
    if (j0 + cc < kocc) Y[(int64_t)i * n + j0 + cc] = Zb[idx];
  }
}

// T factors of the WY blocks of the systems above 256 AOs (the register
// reduction forms its own), one 512-thread block per system after the blocked
// reduction.  BIG (n > kMaxAO): the Gram entries in the system's global slice.
constexpr int kWyGram = kBackChunk * (kBackChunk - 1) / 2 * ((kMaxAO - 2 + kBackChunk - 1) / kBackChunk);
template <bool BIG>
__global__ __launch_bounds__(512) void k_wy_factors(Topo t, const int* order, const double* Aall,
                                                    double* Tall, const unsigned char* active, double* big) {
  const int s = order[blockIdx.x];
  if (active && !active[s]) return;
  const int n = t.ao_off[s + 1] - t.ao_off[s];
  __shared__ double gram_l[BIG ? 1 : kWyGram];
  double* gram = BIG ? big + big_offset(t, s) : gram_l;  // 28 ceil((n-2)/8) <= kBigPerAO n doubles
  wy_factors<1>(Aall + t.mat_off[s], n, Tall + wy_offset(t, s), gram);
}

// n > 256: Y = Q Z on a block of cw columns (cw | 64), the reflectors applied
// kBackChunk at a time in compact WY form with the T factors of k_wy_factors:
//   w = V^T z, y = T w, z -= V y,
// four barriers per block of eight reflectors.  Thread (c, g) = (tid % cw,
// tid / cw) holds column j0 + c, rows i0 + g, i0 + g + 512 / cw, ... of the
// rows the block touches (i0 = one past its lowest reflector); the dot
// partials are summed over the row groups of a wavefront by lane exchanges,
// then over the wavefronts in LDS.
// Dynamic LDS (doubles): non-BIG, the column block [n][cw], the block's
// reflectors [8][n] (zero-padded to absolute rows), partials [8][8][cw] and
// y [8][cw]; BIG (n > kMaxAO), partials and y only: the column block is
// updated in place in Y and the reflectors are read from A.
constexpr int kWyThreads = 512;
__host__ __device__ constexpr size_t backtransform_wy_lds(int n, int cw, bool big) {
  return ((big ? 0 : (size_t)n * cw + (size_t)kBackChunk * n) + (size_t)kBackChunk * 9 * cw) * sizeof(double);
}

template <bool BIG>
__global__ __launch_bounds__(kWyThreads) void k_backtransform_wy(Topo t, const BackTask* tasks,
                                                                 const double* Aall, const double* Zall,
                                                                 const double* Tall, double* Yall,
                                                                 const int* kocc_all, const double* zscale_all,
                                                                 const unsigned char* active) {
  extern __shared__ double lds[];
  const BackTask task = tasks[blockIdx.x];
  const int s = task.sys;
  if (active && !active[s]) return;
  const int n = t.ao_off[s + 1] - t.ao_off[s];
  const int kocc = kocc_all[s];
  if (task.j0 >= kocc) return;
  const int64_t mo = t.mat_off[s];
  const double* __restrict__ A = Aall + mo;
  const double* __restrict__ Z = Zall + mo;
  double* __restrict__ Y = Yall + mo;
  const double* zscale = zscale_all + t.ao_off[s];
  const double* Tf = Tall + wy_offset(t, s);
  const int cw = task.cw, j0 = task.j0;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int c = tid % cw, g = tid / cw, ng = kWyThreads / cw;
  const int col = j0 + c;
  const bool on = col < kocc;
  double* zb = BIG ? Y + j0 : lds;  // z(i, c) at zb[i * ldz + c]
  const int64_t ldz = BIG ? n : cw;
  double* vs = lds + (BIG ? 0 : (int64_t)n * cw);          // [8][n]
  double* part = vs + (BIG ? 0 : (int64_t)kBackChunk * n);  // [8][8][cw]
  double* yv = part + kBackChunk * 8 * cw;                  // [8][cw]
  for (int i = g; i < n; i += ng) {
    const double zv = on ? zscale[col] * Z[(int64_t)i * n + col] : 0.0;
    if (!BIG || on) zb[i * ldz + c] = zv;
  }
  __syncthreads();
  const int nref = n - 2;  // reflectors k = 0 .. n-3
  for (int ch = 0; ch * kBackChunk < nref; ++ch) {
    const int kb = n - 3 - ch * kBackChunk;  // k_0; k_u = kb - u
    const int i0 = max(kb - (kBackChunk - 1), 0) + 1;
    const int m = n - i0;
    if constexpr (!BIG) {
      for (int x = tid; x < kBackChunk * m; x += kWyThreads) {
        const int u = x / m, i = i0 + (x - u * m), k = kb - u;
        vs[u * n + i] = (k >= 0 && i > k) ? A[i + (int64_t)k * n] : 0.0;
      }
      __syncthreads();
    }
    auto vat = [&](int i, int u) __attribute__((always_inline)) -> double {
      if constexpr (BIG) {
        const int k = kb - u;
        return (k >= 0 && i > k) ? A[i + (int64_t)k * n] : 0.0;
      } else {
        return vs[u * n + i];
      }
    };
    double p[kBackChunk];
#pragma unroll
    for (int u = 0; u < kBackChunk; ++u) p[u] = 0.0;
    for (int i = i0 + g; i < n; i += ng) {
      const double zv = (!BIG || on) ? zb[i * ldz + c] : 0.0;
#pragma unroll
      for (int u = 0; u < kBackChunk; ++u) p[u] = fma(vat(i, u), zv, p[u]);
    }
    for (int off = cw; off < 64; off <<= 1)
#pragma unroll
      for (int u = 0; u < kBackChunk; ++u) p[u] += __shfl_xor(p[u], off, 64);
    if (lane < cw)
#pragma unroll
      for (int u = 0; u < kBackChunk; ++u) part[(u * 8 + w) * cw + c] = p[u];
    __syncthreads();
    // y = T w, T lower triangular (thread (u, c))
    if (tid < kBackChunk * cw) {
      const int u = tid / cw, cc = tid - u * cw;
      double acc = 0.0;
      for (int i = 0; i <= u; ++i) {
        double wsum = 0.0;
#pragma unroll
        for (int q = 0; q < 8; ++q) wsum += part[(i * 8 + q) * cw + cc];
        acc = fma(Tf[ch * kBackChunk * kBackChunk + u * kBackChunk + i], wsum, acc);
      }
      yv[u * cw + cc] = acc;
    }
    __syncthreads();
    double yc[kBackChunk];
#pragma unroll
    for (int u = 0; u < kBackChunk; ++u) yc[u] = yv[u * cw + c];
    if (!BIG || on)
      for (int i = i0 + g; i < n; i += ng) {
        double zv = zb[i * ldz + c];
#pragma unroll
        for (int u = 0; u < kBackChunk; ++u) zv = fma(-vat(i, u), yc[u], zv);
        zb[i * ldz + c] = zv;
      }
    __syncthreads();
  }
  if constexpr (!BIG) {
    if (on)
      for (int i = g; i < n; i += ng) Y[(int64_t)i * n + col] = zb[i * cw + c];
  }
}

// two blocks per CU (four wavefronts per SIMD): at most 128 VGPRs.  The host
// sends the column blocks of systems of at most 256 AOs here (the register
// path) and larger ones to k_backtransform_wy; the LDS branch stays as the
// kernel's general form.
__global__ __launch_bounds__(kBackThreads, 4) void k_backtransform(Topo t, const BackTask* tasks,
                                                                const double* Aall,
                                                                const double* Zall,
                                                                const double* Tall, double* Yall,
                                                                const int* kocc_all,
                                                                const double* zscale_all,
                                                                const unsigned char* active) {
  extern __shared__ double lds[];
  __shared__ double ring[2 * kBackChunk * 256];
  __shared__ double part[kBackChunk * kBackThreads];
  __shared__ double wf[kBackChunk * 64];
  const BackTask task = tasks[blockIdx.x];
  const int s = task.sys;
  if (active && !active[s]) return;
  const int n = t.ao_off[s + 1] - t.ao_off[s];
  const int64_t mo = t.mat_off[s];
  const int kocc = kocc_all[s];
  if (task.j0 >= kocc) return;
  const double* zscale = zscale_all + t.ao_off[s];
  if (n <= 256)
    backtransform_registers(Aall + mo, Zall + mo, Tall + wy_offset(t, s),
                            Yall + mo, kocc, zscale, n, task.j0, ring, part, wf);
  else
    backtransform_lds(Aall + mo, Zall + mo, Yall + mo, kocc, zscale, n, task.j0, task.cw, lds);
}

}  // namespace xtb
