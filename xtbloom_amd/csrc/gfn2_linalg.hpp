// Dense FP64 linear algebra for the ragged batch: tiled GEMM over a tile
// queue, per-system Cholesky/inverse and the symmetric eigensolver
// (Householder tridiagonalization + implicit QL + back-transformation).
#pragma once

#include "gfn2_device.hpp"

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
//   K per system: kdim[sys] if given, else n.
//   SYMM: only tiles with ti >= tj are queued; the kernel mirrors them.
template <bool TA, bool TB, bool SYMM>
__global__ __launch_bounds__(256) void k_gemm(Topo t, const TileTask* tasks, const double* A,
                                              const double* Bm, double* C, const double* scale,
                                              const int* kdim, const unsigned char* active) {
  const TileTask task = tasks[blockIdx.x];
  const int s = task.sys;
  if (active && !active[s]) return;
  const int n = t.ao_off[s + 1] - t.ao_off[s];
  const int64_t mo = t.mat_off[s];
  const double* a = A + mo;
  const double* b = Bm + mo;
  double* c = C + mo;
  const double* sc = scale ? scale + t.ao_off[s] : nullptr;
  const int K = kdim ? kdim[s] : n;
  const int i0 = task.ti * kTile, j0 = task.tj * kTile;
  __shared__ double As[kTileK][kTile + 1];
  __shared__ double Bs[kTileK][kTile + 1];
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;  // 16 x 16 threads, 4 x 4 each
  double acc[4][4];
#pragma unroll
  for (int p = 0; p < 4; ++p)
#pragma unroll
    for (int q = 0; q < 4; ++q) acc[p][q] = 0.0;
  for (int k0 = 0; k0 < K; k0 += kTileK) {
    // load A tile: rows i0..i0+63, k k0..k0+15 -> As[k][i]
    for (int idx = threadIdx.x; idx < kTile * kTileK; idx += 256) {
      int kk, ii;
      if (TA) { ii = idx % kTile; kk = idx / kTile; } else { kk = idx % kTileK; ii = idx / kTileK; }
      const int gi = i0 + ii, gk = k0 + kk;
      double val = 0.0;
      if (gi < n && gk < K) {
        val = TA ? a[(int64_t)gk * n + gi] : a[(int64_t)gi * n + gk];
        if (sc) val *= sc[gk];
      }
      As[kk][ii] = val;
    }
    for (int idx = threadIdx.x; idx < kTile * kTileK; idx += 256) {
      int kk, jj;
      if (TB) { kk = idx % kTileK; jj = idx / kTileK; } else { jj = idx % kTile; kk = idx / kTile; }
      const int gj = j0 + jj, gk = k0 + kk;
      double val = 0.0;
      if (gj < n && gk < K) val = TB ? b[(int64_t)gj * n + gk] : b[(int64_t)gk * n + gj];
      Bs[kk][jj] = val;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < kTileK; ++kk) {
      double ar[4], br[4];
#pragma unroll
      for (int p = 0; p < 4; ++p) ar[p] = As[kk][ty + 16 * p];
#pragma unroll
      for (int q = 0; q < 4; ++q) br[q] = Bs[kk][tx + 16 * q];
#pragma unroll
      for (int p = 0; p < 4; ++p)
#pragma unroll
        for (int q = 0; q < 4; ++q) acc[p][q] = fma(ar[p], br[q], acc[p][q]);
    }
    __syncthreads();
  }
#pragma unroll
  for (int p = 0; p < 4; ++p) {
    const int gi = i0 + ty + 16 * p;
    if (gi >= n) continue;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int gj = j0 + tx + 16 * q;
      if (gj >= n) continue;
      c[(int64_t)gi * n + gj] = acc[p][q];
      if (SYMM && task.ti != task.tj) c[(int64_t)gj * n + gi] = acc[p][q];
    }
  }
}

// Per-system Cholesky S = L L^T and X = L^{-T} (row-major; X upper triangular).
// L is built in `work` (column-major lower).  status[s] = 8 on a non-positive
// pivot or when the conditioning bound 1/(|S|_1 |L^-1|_1 |L^-T|_1) < 1e-12.
__global__ __launch_bounds__(256) void k_cholesky_inverse(Topo t, const int* order, const double* S,
                                                          double* work, double* X, SysState* status) {
  const int s = order[blockIdx.x];
  const int n = t.ao_off[s + 1] - t.ao_off[s];
  const int64_t mo = t.mat_off[s];
  const double* Sm = S + mo;
  double* L = work + mo;  // column-major: L[i + j*n]
  double* Xm = X + mo;
  __shared__ double red[16];
  __shared__ int bad;
  if (threadIdx.x == 0) bad = 0;
  for (int64_t idx = threadIdx.x; idx < (int64_t)n * n; idx += blockDim.x) L[idx] = Sm[idx];
  __syncthreads();
  // |S|_1
  double colmax = 0.0;
  for (int j = threadIdx.x; j < n; j += blockDim.x) {
    double sacc = 0.0;
    for (int i = 0; i < n; ++i) sacc += fabs(Sm[(int64_t)i * n + j]);
    colmax = fmax(colmax, sacc);
  }
  // block max via shared array
  __shared__ double mx[256];
  mx[threadIdx.x] = colmax;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) mx[threadIdx.x] = fmax(mx[threadIdx.x], mx[threadIdx.x + w]);
    __syncthreads();
  }
  const double snorm = mx[0];
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
  // conditioning bound: |S^-1|_1 <= |Linv^T|_1 |Linv|_1 = (max row-sum of X^T...) use both norms of X
  double c1 = 0.0, ci = 0.0;
  for (int j = threadIdx.x; j < n; j += blockDim.x) {
    double rs = 0.0, cs = 0.0;
    for (int i = 0; i < n; ++i) {
      rs += fabs(Xm[(int64_t)j * n + i]);
      cs += fabs(Xm[(int64_t)i * n + j]);
    }
    c1 = fmax(c1, rs);
    ci = fmax(ci, cs);
  }
  mx[threadIdx.x] = c1;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) mx[threadIdx.x] = fmax(mx[threadIdx.x], mx[threadIdx.x + w]);
    __syncthreads();
  }
  const double n1 = mx[0];
  __syncthreads();
  mx[threadIdx.x] = ci;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) mx[threadIdx.x] = fmax(mx[threadIdx.x], mx[threadIdx.x + w]);
    __syncthreads();
  }
  const double n2 = mx[0];
  if (threadIdx.x == 0) {
    const double rcond_bound = 1.0 / (snorm * n1 * n2);
    if (bad || !(rcond_bound >= 1e-12) || !isfinite(rcond_bound)) status[s].status = 8;
  }
  (void)red;
}

// ---------------------------------------------------------------------------
// Symmetric eigensolver, one 256-thread block per system.
//   A (n x n, symmetric, row-major == column-major) is destroyed: its strict
//   lower triangle receives the Householder vectors.  Z receives the
//   eigenvectors (column-major: vector j is Z[j*n .. j*n+n)), w the ascending
//   eigenvalues.  status[s] = 8 when QL fails to converge.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_eigh(Topo t, const int* order, double* Aall, double* Zall,
                                              double* Yall, double* wall, SysState* status,
                                              const unsigned char* active) {
  const int s = order[blockIdx.x];
  if (active && !active[s]) return;
  const int n = t.ao_off[s + 1] - t.ao_off[s];
  const int64_t mo = t.mat_off[s];
  double* A = Aall + mo;  // column-major view
  double* Z = Zall + mo;  // column-major eigenvectors of T
  double* Y = Yall + mo;  // column-major eigenvectors of A (sorted)
  double* w = wall + t.ao_off[s];
  __shared__ double v[kMaxAO];
  __shared__ double pq[kMaxAO];
  __shared__ double d[kMaxAO];
  __shared__ double e[kMaxAO];
  __shared__ double red[16];
  __shared__ double bc[2];
  __shared__ int ctrl[4];
  const int tid = threadIdx.x, nt = blockDim.x;

  // ---- Householder tridiagonalization ----
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
        // already reduced: no reflection
        e[k] = x0;
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
    // p = A22 v
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
    ctrl[3] = 0;
  }
  // Z = I
  for (int64_t idx = tid; idx < (int64_t)n * n; idx += nt) Z[idx] = (idx % (n + 1) == 0) ? 1.0 : 0.0;
  __syncthreads();

  // ---- implicit QL with Wilkinson-type shift; rotations applied to Z rows ----
  // pq[] holds cosines, v[] holds sines of the current sweep.
  for (int l = 0; l < n; ++l) {
    int iter = 0;
    for (;;) {
      if (tid == 0) {
        int m = l;
        for (; m < n - 1; ++m) {
          const double dd = fabs(d[m]) + fabs(d[m + 1]);
          if (fabs(e[m]) <= 2.220446049250313e-16 * dd) break;
        }
        ctrl[0] = m;
        ctrl[1] = m;  // last rotation index applied is ctrl[1]..m-1 -> set below
        if (m != l) {
          if (++iter > 60) {
            ctrl[3] = 1;
            ctrl[0] = l;  // give up on this eigenvalue
          } else {
            double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
            double r = hypot(g, 1.0);
            g = d[m] - d[l] + e[l] / (g + (g >= 0.0 ? fabs(r) : -fabs(r)));
            double sn = 1.0, cs = 1.0, p = 0.0;
            int i = m - 1;
            bool underflow = false;
            for (; i >= l; --i) {
              double f = sn * e[i], b = cs * e[i];
              r = hypot(f, g);
              e[i + 1] = r;
              if (r == 0.0) {
                d[i + 1] -= p;
                e[m] = 0.0;
                underflow = true;
                break;
              }
              sn = f / r;
              cs = g / r;
              g = d[i + 1] - p;
              r = (d[i] - g) * sn + 2.0 * cs * b;
              p = sn * r;
              d[i + 1] = g + p;
              g = cs * r - b;
              pq[i] = cs;
              v[i] = sn;
            }
            ctrl[1] = underflow ? i + 1 : l;  // rotations i in [ctrl1, m-1]
            if (!underflow) {
              d[l] -= p;
              e[l] = g;
              e[m] = 0.0;
            }
          }
        }
      }
      __syncthreads();
      const int m = ctrl[0], lo = ctrl[1];
      if (m == l) break;
      // apply rotations i = m-1 .. lo to columns (i, i+1) of Z, row by row
      for (int r = tid; r < n; r += nt) {
        double f = Z[r + (int64_t)m * n];
        for (int i = m - 1; i >= lo; --i) {
          const double zi = Z[r + (int64_t)i * n];
          Z[r + (int64_t)(i + 1) * n] = v[i] * zi + pq[i] * f;
          f = pq[i] * zi - v[i] * f;
        }
        Z[r + (int64_t)lo * n] = f;
      }
      __syncthreads();
    }
    __syncthreads();
  }
  if (tid == 0 && ctrl[3]) status[s].status = 8;

  // ---- back-transformation Z <- H_0 ... H_{n-3} Z (waves over columns) ----
  const int lane = tid & 63, wid = tid >> 6, nw = nt >> 6;
  for (int k = n - 3; k >= 0; --k) {
    const int m = n - k - 1;
    const double* hv = A + (k + 1) + (int64_t)k * n;
    for (int j = wid; j < n; j += nw) {
      double* zc = Z + (int64_t)j * n + (k + 1);
      double dot = 0.0;
      for (int i = lane; i < m; i += 64) dot += hv[i] * zc[i];
      dot = wave_sum(dot);
      for (int i = lane; i < m; i += 64) zc[i] -= 2.0 * dot * hv[i];
    }
    __syncthreads();
  }
  // ---- ascending order: rank sort, stable on ties ----
  for (int j = tid; j < n; j += nt) {
    const double dj = d[j];
    int rank = 0;
    for (int i = 0; i < n; ++i) {
      const double di = d[i];
      rank += (di < dj) || (di == dj && i < j);
    }
    pq[j] = (double)rank;
  }
  __syncthreads();
  for (int j = 0; j < n; ++j) {
    const int rk = (int)pq[j];
    for (int i = tid; i < n; i += nt) Y[(int64_t)rk * n + i] = Z[(int64_t)j * n + i];
    if (tid == 0) w[rk] = d[j];
  }
}

}  // namespace xtb
