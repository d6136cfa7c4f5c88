// Shell-pair overlap / dipole / traceless-quadrupole integrals for s, p, d shells.
//
// One thread owns one shell pair <bra| O |ket> with the operator origin on the
// ket atom (reference src/model/common/integrals.cpp:387-691 and
// gfn2/integrals.hpp:25-45).  Angular momenta are template parameters so every
// (l_bra, l_ket) combination is a separate fully unrolled instantiation that
// keeps its Cartesian blocks in registers.
//
// Two entry points:
//   pair_integrals   -> spherical S, D(3), Q(6) blocks of one ordered shell pair
//   pair_gradient    -> sum_{blocks} adjoint . d(block)/dR_ket, contracted in the
//                       Cartesian frame inside the primitive loop; no derivative
//                       block is ever materialized.
#pragma once

#include "gfn2_device.hpp"

namespace xtb {

template <int L> struct Cart;
template <> struct Cart<0> {
  static constexpr int n = 1;
  __device__ static constexpr int e(int, int) { return 0; }
};
template <> struct Cart<1> {
  static constexpr int n = 3;
  __device__ static constexpr int e(int i, int x) { return i == x ? 1 : 0; }
};
template <> struct Cart<2> {
  static constexpr int n = 6;
  // xx, xy, xz, yy, yz, zz
  __device__ static constexpr int e(int i, int x) {
    return i == 0 ? (x == 0 ? 2 : 0)
         : i == 1 ? (x == 2 ? 0 : 1)
         : i == 2 ? (x == 1 ? 0 : 1)
         : i == 3 ? (x == 1 ? 2 : 0)
         : i == 4 ? (x == 0 ? 0 : 1)
                  : (x == 2 ? 2 : 0);
  }
};

// operator powers: x y z xx xy yy xz yz zz
__device__ __forceinline__ constexpr int moment_pow(int c, int x) {
  return c == 0 ? (x == 0) : c == 1 ? (x == 1) : c == 2 ? (x == 2)
       : c == 3 ? (x == 0 ? 2 : 0) : c == 4 ? (x == 2 ? 0 : 1) : c == 5 ? (x == 1 ? 2 : 0)
       : c == 6 ? (x == 1 ? 0 : 1) : c == 7 ? (x == 0 ? 0 : 1) : (x == 2 ? 2 : 0);
}

// real-spherical transform coefficient T[sph][cart] (standard order, p = [y, z, x])
template <int L> __device__ __forceinline__ constexpr double tsph(int s, int c);
template <> __device__ __forceinline__ constexpr double tsph<0>(int, int) { return 1.0; }
template <> __device__ __forceinline__ constexpr double tsph<1>(int s, int c) {
  return (s == 0 && c == 1) || (s == 1 && c == 2) || (s == 2 && c == 0) ? 1.0 : 0.0;
}
template <> __device__ __forceinline__ constexpr double tsph<2>(int s, int c) {
  constexpr double r3 = 1.732050807568877293527446341505872367;
  return s == 0 ? (c == 1 ? r3 : 0.0)
       : s == 1 ? (c == 4 ? r3 : 0.0)
       : s == 2 ? (c == 0 || c == 3 ? -0.5 : (c == 5 ? 1.0 : 0.0))
       : s == 3 ? (c == 2 ? r3 : 0.0)
                : (c == 0 ? 0.5 * r3 : (c == 3 ? -0.5 * r3 : 0.0));
}

// 1D Obara-Saika overlap table o[a][b], a about the ket, b about the bra.
template <int AMAX, int BMAX>
__device__ __forceinline__ void os1d(double pK, double pB, double h, double (&o)[AMAX + 1][BMAX + 1]) {
  o[0][0] = 1.0;
#pragma unroll
  for (int a = 0; a < AMAX; ++a) o[a + 1][0] = pK * o[a][0] + (a > 0 ? a * h * o[a - 1][0] : 0.0);
#pragma unroll
  for (int b = 0; b < BMAX; ++b)
#pragma unroll
    for (int a = 0; a <= AMAX; ++a)
      o[a][b + 1] = pB * o[a][b] + h * ((a > 0 ? a * o[a - 1][b] : 0.0) + (b > 0 ? b * o[a][b - 1] : 0.0));
}

template <int LB, int LK>
struct PairBlock {
  static constexpr int nb = 2 * LB + 1, nk = 2 * LK + 1;
  double s[nb][nk];
  double m[9][nb][nk];  // dipole x,y,z then traceless xx,xy,yy,xz,yz,zz
};

// Spherical S and (optionally) D/Q of bra shell `bra`, ket shell `ket`; v = R_ket - R_bra.
template <int LB, int LK, bool MOM>
__device__ void pair_integrals(const Topo& t, int bra, int ket, const double v[3],
                               PairBlock<LB, LK>& out) {
  constexpr int CB = Cart<LB>::n, CK = Cart<LK>::n;
  constexpr int AMAX = LK + (MOM ? 2 : 0);
  double cs[CB][CK];
  double cm[MOM ? 9 : 1][CB][CK];
#pragma unroll
  for (int i = 0; i < CB; ++i)
#pragma unroll
    for (int j = 0; j < CK; ++j) {
      cs[i][j] = 0.0;
#pragma unroll
      for (int c = 0; c < (MOM ? 9 : 1); ++c) cm[c][i][j] = 0.0;
    }
  const double r2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
  const int pk0 = t.sh_prim[ket], npk = t.sh_nprim[ket];
  const int pb0 = t.sh_prim[bra], npb = t.sh_nprim[bra];
  for (int ik = 0; ik < npk; ++ik) {
    const double ak = t.alpha[pk0 + ik], ck = t.coef[pk0 + ik];
    for (int ib = 0; ib < npb; ++ib) {
      const double ab = t.alpha[pb0 + ib];
      const double p = ak + ab, ip = 1.0 / p;
      const double est = ak * ab * r2 * ip;
      if (est > 25.0) continue;
      const double sq = sqrt(kPi * ip);
      const double pref = exp(-est) * sq * sq * sq * ck * t.coef[pb0 + ib];
      double o[3][AMAX + 1][LB + 1];
#pragma unroll
      for (int x = 0; x < 3; ++x) os1d<AMAX, LB>(-v[x] * ab * ip, v[x] * ak * ip, 0.5 * ip, o[x]);
#pragma unroll
      for (int i = 0; i < CB; ++i)
#pragma unroll
        for (int j = 0; j < CK; ++j) {
          const int bx = Cart<LB>::e(i, 0), by = Cart<LB>::e(i, 1), bz = Cart<LB>::e(i, 2);
          const int kx = Cart<LK>::e(j, 0), ky = Cart<LK>::e(j, 1), kz = Cart<LK>::e(j, 2);
          cs[i][j] += pref * o[0][kx][bx] * o[1][ky][by] * o[2][kz][bz];
          if constexpr (MOM) {
#pragma unroll
            for (int c = 0; c < 9; ++c)
              cm[c][i][j] += pref * o[0][kx + moment_pow(c, 0)][bx] *
                             o[1][ky + moment_pow(c, 1)][by] * o[2][kz + moment_pow(c, 2)][bz];
          }
        }
    }
  }
  // spherical transform
#pragma unroll
  for (int u = 0; u < 2 * LB + 1; ++u)
#pragma unroll
    for (int w = 0; w < 2 * LK + 1; ++w) {
      double vs = 0.0, vm[MOM ? 9 : 1];
#pragma unroll
      for (int c = 0; c < (MOM ? 9 : 1); ++c) vm[c] = 0.0;
#pragma unroll
      for (int i = 0; i < CB; ++i) {
        const double tb = tsph<LB>(u, i);
        if (tb == 0.0) continue;
#pragma unroll
        for (int j = 0; j < CK; ++j) {
          const double tk = tsph<LK>(w, j);
          if (tk == 0.0) continue;
          vs += tb * cs[i][j] * tk;
          if constexpr (MOM) {
#pragma unroll
            for (int c = 0; c < 9; ++c) vm[c] += tb * cm[c][i][j] * tk;
          }
        }
      }
      out.s[u][w] = vs;
      if constexpr (MOM) {
        const double tr = 0.5 * (vm[3] + vm[5] + vm[8]);
        out.m[0][u][w] = vm[0];
        out.m[1][u][w] = vm[1];
        out.m[2][u][w] = vm[2];
        out.m[3][u][w] = 1.5 * vm[3] - tr;
        out.m[4][u][w] = 1.5 * vm[4];
        out.m[5][u][w] = 1.5 * vm[5] - tr;
        out.m[6][u][w] = 1.5 * vm[6];
        out.m[7][u][w] = 1.5 * vm[7];
        out.m[8][u][w] = 1.5 * vm[8] - tr;
      }
    }
}

// Spherical adjoints of one ordered block: a[0] for S, a[1..3] for D, a[4..9] for
// traceless Q ([xx,xy,yy,xz,yz,zz]).
template <int LB, int LK>
struct PairAdjoint {
  double a[10][2 * LB + 1][2 * LK + 1];
};

// g += sum adjoint . d(block)/dR_ket (operator origin moving with the ket).
template <int LB, int LK>
__device__ void pair_gradient(const Topo& t, int bra, int ket, const double v[3],
                              const PairAdjoint<LB, LK>& adj, double g[3]) {
  constexpr int CB = Cart<LB>::n, CK = Cart<LK>::n;
  constexpr int AMAX = LK + 3;
  // pull the spherical adjoints back to the Cartesian raw-moment frame
  double ca[10][CB][CK];
#pragma unroll
  for (int i = 0; i < CB; ++i)
#pragma unroll
    for (int j = 0; j < CK; ++j) {
      double raw[10];
#pragma unroll
      for (int c = 0; c < 10; ++c) raw[c] = 0.0;
#pragma unroll
      for (int u = 0; u < 2 * LB + 1; ++u) {
        const double tb = tsph<LB>(u, i);
        if (tb == 0.0) continue;
#pragma unroll
        for (int w = 0; w < 2 * LK + 1; ++w) {
          const double tk = tsph<LK>(w, j);
          if (tk == 0.0) continue;
          const double f = tb * tk;
#pragma unroll
          for (int c = 0; c < 10; ++c) raw[c] += f * adj.a[c][u][w];
        }
      }
      // traceless projection Q = 1.5 M - 0.5 tr(M) I pulled back to raw moments
      const double diag = raw[4] + raw[6] + raw[9];
      ca[0][i][j] = raw[0];
      ca[1][i][j] = raw[1];
      ca[2][i][j] = raw[2];
      ca[3][i][j] = raw[3];
      ca[4][i][j] = 1.5 * raw[4] - 0.5 * diag;  // xx
      ca[5][i][j] = 1.5 * raw[5];               // xy
      ca[6][i][j] = 1.5 * raw[6] - 0.5 * diag;  // yy
      ca[7][i][j] = 1.5 * raw[7];               // xz
      ca[8][i][j] = 1.5 * raw[8];               // yz
      ca[9][i][j] = 1.5 * raw[9] - 0.5 * diag;  // zz
    }
  const double r2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
  const int pk0 = t.sh_prim[ket], npk = t.sh_nprim[ket];
  const int pb0 = t.sh_prim[bra], npb = t.sh_nprim[bra];
  double gx = 0.0, gy = 0.0, gz = 0.0;
  for (int ik = 0; ik < npk; ++ik) {
    const double ak = t.alpha[pk0 + ik], ck = t.coef[pk0 + ik];
    for (int ib = 0; ib < npb; ++ib) {
      const double ab = t.alpha[pb0 + ib];
      const double p = ak + ab, ip = 1.0 / p;
      const double est = ak * ab * r2 * ip;
      if (est > 25.0) continue;
      const double sq = sqrt(kPi * ip);
      const double pref = exp(-est) * sq * sq * sq * ck * t.coef[pb0 + ib];
      double o[3][AMAX + 1][LB + 1];
#pragma unroll
      for (int x = 0; x < 3; ++x) os1d<AMAX, LB>(-v[x] * ab * ip, v[x] * ak * ip, 0.5 * ip, o[x]);
      double px = 0.0, py = 0.0, pz = 0.0;
#pragma unroll
      for (int i = 0; i < CB; ++i)
#pragma unroll
        for (int j = 0; j < CK; ++j) {
          const int bx = Cart<LB>::e(i, 0), by = Cart<LB>::e(i, 1), bz = Cart<LB>::e(i, 2);
#pragma unroll
          for (int c = 0; c < 10; ++c) {
            // c == 0: overlap (no operator power); c >= 1: moment c-1
            const int mx = c == 0 ? 0 : moment_pow(c - 1, 0);
            const int my = c == 0 ? 0 : moment_pow(c - 1, 1);
            const int mz = c == 0 ? 0 : moment_pow(c - 1, 2);
            const int ax = Cart<LK>::e(j, 0) + mx, ay = Cart<LK>::e(j, 1) + my,
                      az = Cart<LK>::e(j, 2) + mz;
            const double fx = o[0][ax][bx], fy = o[1][ay][by], fz = o[2][az][bz];
            const double dx = 2.0 * ak * o[0][ax + 1][bx] - (ax > 0 ? ax * o[0][ax - 1][bx] : 0.0);
            const double dy = 2.0 * ak * o[1][ay + 1][by] - (ay > 0 ? ay * o[1][ay - 1][by] : 0.0);
            const double dz = 2.0 * ak * o[2][az + 1][bz] - (az > 0 ? az * o[2][az - 1][bz] : 0.0);
            const double w = ca[c][i][j];
            px += w * dx * fy * fz;
            py += w * fx * dy * fz;
            pz += w * fx * fy * dz;
          }
        }
      gx += pref * px;
      gy += pref * py;
      gz += pref * pz;
    }
  }
  g[0] += gx;
  g[1] += gy;
  g[2] += gz;
}

}  // namespace xtb
