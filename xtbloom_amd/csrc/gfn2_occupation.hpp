// Fermi occupations: the binary64 policy of the reference
// (gfn2/occupation_binary64_policy.hpp, used by its CUDA backend always and
// by its CPU path when the long-double root's rounded occupations fail to
// conserve the count, eigensolver.cpp:1162-1218).  Here it is the rare path
// behind the wave-parallel FP64 bisection of the eigen kernel: one lane runs
// it when the published occupations of the fast path miss the count by more
// than 64 eps (degenerate frontiers whose count has no equal binary64 split).
//   root        bisection in the frame translated to the lowest (electrons) or
//               highest (holes) level and scaled by kT, saturated arithmetic,
//               compensated sums; one retry at the unique changing degenerate
//               frontier when the binary64 spacing is exhausted
//   publication at most two degenerate blocks take one shared occupation each
//               (candidates around the count they should hold), chosen by the
//               smallest compensated error so that the double-double audited
//               count is within 64 eps (strict) or 2 eps per block member
//               (relaxed, degenerate blocks only)
#pragma once

namespace xtb {

constexpr double P64_MAX = 1.79769313486231570814527423731704357e308;
constexpr double P64_EPS = 2.220446049250313080847263336181640625e-16;

__device__ inline double p64_sat_add(double a, double b) {
  const double r = a + b;
  if (isfinite(r)) return r;
  return (signbit(a) != 0) == (signbit(b) != 0) && signbit(a) ? -P64_MAX : P64_MAX;
}
__device__ inline double p64_sat_sub(double a, double b) {
  const double r = a - b;
  if (isfinite(r)) return r;
  return a < b ? -P64_MAX : P64_MAX;
}
__device__ inline double p64_sat_mul_nonneg(double a, double b) {
  if (a == 0.0 || b == 0.0) return 0.0;
  return a > P64_MAX / b ? P64_MAX : a * b;
}
__device__ inline double p64_sat_affine(double ref, double mult, double scale) {
  const double prod = mult * scale;
  if (isfinite(prod)) return p64_sat_add(ref, prod);
  const double nrm = ref / P64_MAX + mult * (scale / P64_MAX);
  if (nrm >= 1.0) return P64_MAX;
  if (nrm <= -1.0) return -P64_MAX;
  return nrm * P64_MAX;
}
// (e - ref) / T without spurious overflow
__device__ inline double p64_scaled(double e, double ref, double t) {
  double r;
  if ((signbit(e) != 0) == (signbit(ref) != 0)) r = (e - ref) / t;
  else r = e / t - ref / t;
  if (isfinite(r)) return r;
  return e < ref ? -P64_MAX : P64_MAX;
}
__device__ inline double p64_fermi(double se, double smu) {
  const double a = p64_sat_sub(se, smu);
  if (a >= 0.0) { const double v = exp(-a); return v / (1.0 + v); }
  return 1.0 / (exp(a) + 1.0);
}
__device__ inline double p64_hole(double se, double smu) {
  const double a = p64_sat_sub(se, smu);
  if (a >= 0.0) return 1.0 / (exp(-a) + 1.0);
  const double v = exp(a);
  return v / (1.0 + v);
}
struct p64_ksum { double v, c; };  // Kahan sum and its compensation
__device__ inline void p64_kadd(p64_ksum* s, double x) {
  const double y = x - s->c, t = s->v + y;
  s->c = (t - s->v) - y;
  s->v = t;
}
__device__ inline double p64_quantity(const double* e, int n, double ref, double smu, double t, int holes) {
  p64_ksum s = {0.0, 0.0};
  for (int i = 0; i < n; ++i) {
    const double se = p64_scaled(e[i], ref, t);
    p64_kadd(&s, holes ? p64_hole(se, smu) : p64_fermi(se, smu));
  }
  return s.v;
}
struct p64_root {
  double ref, lo, hi, smu, q;
  int exhausted, retried, fb, fe;
};
__device__ inline int p64_root_once(const double* e, int n, double target, double t, int holes, double ref,
                         p64_root* r) {
  const double lf = log(target) - log((double)n);
  const double steps = fmax(64.0, -lf + 8.0);
  const double smin = p64_scaled(e[0], ref, t), smax = p64_scaled(e[n - 1], ref, t);
  const double scale = fmax(1.0, fabs(p64_sat_sub(smax, smin)));
  const double margin = p64_sat_add(steps, p64_sat_mul_nonneg(64.0 * P64_EPS, scale));
  double lo = p64_sat_sub(smin, margin), hi = p64_sat_add(smax, margin);
  const double qlo = p64_quantity(e, n, ref, lo, t, holes), qhi = p64_quantity(e, n, ref, hi, t, holes);
  const int brk = holes ? (qlo >= target && qhi <= target) : (qlo <= target && qhi >= target);
  if (!isfinite(lo) || !isfinite(hi) || !(lo < hi) || !isfinite(qlo) || !isfinite(qhi) || !brk) return 0;
  const double tol = 64.0 * P64_EPS * target;
  int exhausted = 0;
  for (int it = 0; it < 4096; ++it) {
    const double mid = 0.5 * lo + 0.5 * hi;
    const double q = p64_quantity(e, n, ref, mid, t, holes);
    if (!isfinite(q)) return 0;
    if (fabs(q - target) <= tol) { lo = hi = mid; break; }
    if (mid == lo || mid == hi) { exhausted = 1; break; }
    if ((!holes && q < target) || (holes && q > target)) lo = mid; else hi = mid;
  }
  const double smu = 0.5 * lo + 0.5 * hi;
  const double q = p64_quantity(e, n, ref, smu, t, holes);
  if (!isfinite(q)) return 0;
  r->ref = ref; r->lo = lo; r->hi = hi; r->smu = smu; r->q = q;
  r->exhausted = exhausted; r->retried = 0; r->fb = -1; r->fe = -1;
  return 1;
}
// the one degenerate block whose occupation differs between the bracket ends
// (n: none, -1: more than one)
__device__ inline int p64_frontier(const double* e, int n, double t, int holes, const p64_root* r) {
  int fr = n;
  for (int b = 0; b < n;) {
    int en = b + 1;
    while (en < n && e[en] == e[b]) ++en;
    if (en - b > 1) {
      const double se = p64_scaled(e[b], r->ref, t);
      const double lv = holes ? p64_hole(se, r->lo) : p64_fermi(se, r->lo);
      const double uv = holes ? p64_hole(se, r->hi) : p64_fermi(se, r->hi);
      if (lv != uv) {
        if (fr != n) return -1;
        fr = b;
      }
    }
    b = en;
  }
  return fr;
}
__device__ inline int p64_solve_root(const double* e, int n, double target, double t, int holes, p64_root* r) {
  if (!p64_root_once(e, n, target, t, holes, holes ? e[n - 1] : e[0], r)) return 0;
  const double tol = 64.0 * P64_EPS * target;
  const double qlo = p64_quantity(e, n, r->ref, r->lo, t, holes), qhi = p64_quantity(e, n, r->ref, r->hi, t, holes);
  const int straddles = isfinite(qlo) && isfinite(qhi) &&
                        (holes ? (qlo >= target && qhi <= target) : (qlo <= target && qhi >= target));
  if (!r->exhausted || fabs(r->q - target) <= tol || !straddles) return 1;
  const int fr = p64_frontier(e, n, t, holes, r);
  if (fr < 0 || fr >= n) return 1;
  r->fb = fr;
  r->fe = fr + 1;
  while (r->fe < n && e[r->fe] == e[fr]) ++r->fe;
  if (e[fr] == r->ref) return 1;
  p64_root again;
  if (!p64_root_once(e, n, target, t, holes, e[fr], &again)) return 1;
  if (fabs(again.q - target) < fabs(r->q - target)) {
    again.retried = 1;
    again.fb = r->fb;
    again.fe = r->fe;
    *r = again;
  }
  return 1;
}
__device__ inline int p64_largest_block(const double* e, int n) {
  int big = 0;
  for (int b = 0; b < n;) {
    int en = b + 1;
    while (en < n && e[en] == e[b]) ++en;
    if (en - b > 1 && en - b > big) big = en - b;
    b = en;
  }
  return big;
}
struct p64_fix { int b, e; double occ; };
struct p64_pub { p64_fix fix[2]; int nfix, relaxed; double rtol, cerr; };
__device__ inline double p64_published(const double* e, int i, double t, const p64_root* r, const p64_pub* p) {
  for (int c = p->nfix; c > 0; --c)
    if (i >= p->fix[c - 1].b && i < p->fix[c - 1].e) return p->fix[c - 1].occ;
  return p64_fermi(p64_scaled(e[i], r->ref, t), r->smu);
}
__device__ inline double p64_pub_error(const double* e, int n, double t, const p64_root* r, const p64_pub* p,
                            int holes, double target) {
  p64_ksum s = {0.0, 0.0};
  for (int i = 0; i < n; ++i) {
    const double o = p64_published(e, i, t, r, p);
    p64_kadd(&s, holes ? 1.0 - o : o);
  }
  return fabs((s.v - target) - s.c);
}
struct p64_dd { double hi, lo; };  // double-double accumulator
__device__ inline void p64_dd_add(p64_dd* d, double x) {
  const double s = d->hi + x, v = s - d->hi;
  const double err = (d->hi - (s - v)) + (x - v) + d->lo;
  const double h = s + err;
  d->lo = err - (h - s);
  d->hi = h;
}
__device__ inline int p64_dd_sign(p64_dd d) {
  if (d.hi > 0.0) return 1;
  if (d.hi < 0.0) return -1;
  return d.lo > 0.0 ? 1 : (d.lo < 0.0 ? -1 : 0);
}
__device__ inline int p64_within(const double* e, int n, double t, const p64_root* r, const p64_pub* p, int holes,
                      double target, double tol) {
  p64_dd d = {0.0, 0.0};
  for (int i = 0; i < n; ++i) {
    const double o = p64_published(e, i, t, r, p);
    p64_dd_add(&d, holes ? 1.0 - o : o);
  }
  p64_dd_add(&d, -target);
  p64_dd below = d, above = d;
  p64_dd_add(&below, -tol);
  p64_dd_add(&above, tol);
  return p64_dd_sign(below) <= 0 && p64_dd_sign(above) >= 0;
}
__device__ inline double p64_entropy(const double* e, int n, double t, const p64_root* r, const p64_pub* p) {
  p64_ksum s = {0.0, 0.0};
  for (int i = 0; i < n; ++i) {
    const double o = p64_published(e, i, t, r, p);
    if (o > 0.0 && o < 1.0) {
      const double h = 1.0 - o;
      p64_kadd(&s, -o * log(o) - h * log(h));
    }
  }
  return s.v;
}
struct p64_cand { int found, frac, b, e; double err, occ, tol; };
__device__ inline int p64_better(double err, int frac, int b, double occ, const p64_cand* best) {
  return !best->found || err < best->err || (err == best->err && frac && !best->frac) ||
         (err == best->err && frac == best->frac && occ < best->occ) ||
         (err == best->err && frac == best->frac && occ == best->occ && b < best->b);
}
__device__ inline int p64_select(const double* e, int n, double target, double t, int holes, const p64_root* r,
                      p64_pub* pub) {
  *pub = p64_pub{};
  const double strict = 64.0 * P64_EPS * target;
  pub->cerr = p64_pub_error(e, n, t, r, pub, holes, target);
  if (p64_within(e, n, t, r, pub, holes, target, strict)) return 1;
  for (int phase = 0; phase < 2; ++phase) {
    p64_cand sc{}, rc{};
    for (int b = 0; b < n;) {
      int en = b + 1;
      while (en < n && e[en] == e[b]) ++en;
      const int cnt = en - b;
      const double old = p64_published(e, b, t, r, pub);
      p64_ksum other = {0.0, 0.0};
      for (int i = 0; i < n; ++i) {
        if (i >= b && i < en) continue;
        const double o = p64_published(e, i, t, r, pub);
        p64_kadd(&other, holes ? 1.0 - o : o);
      }
      const double want = ((target - other.v) + other.c) / (double)cnt;
      const double near = holes ? 1.0 - want : want;
      const double nd = nextafter(near, -P64_MAX), nu = nextafter(near, P64_MAX);
      const double cands[8] = {near, nd, nu, nextafter(nd, -P64_MAX), nextafter(nu, P64_MAX),
                               old, nextafter(old, -P64_MAX), nextafter(old, P64_MAX)};
      for (int ci = 0; ci < 8; ++ci) {
        const double occ = cands[ci];
        int dup = 0;
        for (int pj = 0; pj < ci; ++pj) dup = dup || occ == cands[pj];
        if (!isfinite(occ) || occ < 0.0 || occ > 1.0 || dup) continue;
        if (pub->nfix >= 2) continue;
        p64_pub trial = *pub;
        trial.fix[trial.nfix].b = b;
        trial.fix[trial.nfix].e = en;
        trial.fix[trial.nfix].occ = occ;
        trial.nfix++;
        const double err = p64_pub_error(e, n, t, r, &trial, holes, target);
        const int frac = occ > 0.0 && occ < 1.0;
        if (p64_within(e, n, t, r, &trial, holes, target, strict) && p64_better(err, frac, b, occ, &sc)) {
          sc.found = 1; sc.err = err; sc.frac = frac; sc.b = b; sc.e = en; sc.occ = occ; sc.tol = strict;
        }
        if (phase == 0 && cnt > 1) {
          const double rt = 2.0 * P64_EPS * (double)cnt;
          if (p64_within(e, n, t, r, &trial, holes, target, rt) && p64_better(err, frac, b, occ, &rc)) {
            rc.found = 1; rc.err = err; rc.frac = frac; rc.b = b; rc.e = en; rc.occ = occ; rc.tol = rt;
          }
        }
      }
      b = en;
    }
    const p64_cand* sel = sc.found ? &sc : &rc;
    if (!sel->found)
      return pub->relaxed && p64_within(e, n, t, r, pub, holes, target, pub->rtol);
    pub->fix[pub->nfix].b = sel->b;
    pub->fix[pub->nfix].e = sel->e;
    pub->fix[pub->nfix].occ = sel->occ;
    pub->nfix++;
    pub->cerr = sel->err;
    if (sc.found) {
      pub->relaxed = 0;
      pub->rtol = 0.0;
      break;
    }
    pub->relaxed = 1;
    pub->rtol = sel->tol;
  }
  return p64_within(e, n, t, r, pub, holes, target, strict) ||
         (pub->relaxed && p64_within(e, n, t, r, pub, holes, target, pub->rtol));
}


// The rare path for one spin channel, as the reference's GPU backend runs it
// for every channel (src/backends/cuda/gfn2_occupations.cu:80-152): the
// binary64 root, the publication and the root acceptance at 1024 eps (or the
// frontier floor).  Returns false on a data failure (the system fails with
// EIGENSOLVER_FAILED).  The reference's CPU path enters the policy only
// after its long-double root and then also fails a missed root without a
// degenerate block (eigensolver.cpp:1170-1175); without long double the
// binary64 root is the authoritative one here, as on the reference's GPU.
__device__ inline bool fermi_policy(const double* e, int n, double nel, double kt, double* occ,
                                    double& entropy) {
  const bool holes = nel > 0.5 * n;
  const double cap = (double)n, t64 = holes ? cap - nel : nel;
  p64_root root;
  p64_pub pub;
  if (!(t64 > 0.0) || !isfinite(t64) || !p64_solve_root(e, n, t64, kt, holes, &root)) return false;
  if (!p64_select(e, n, t64, kt, holes, &root, &pub)) return false;
  double rtol = 1024.0 * P64_EPS * t64;
  if (root.fb >= 0 && root.fe > root.fb) rtol = fmax(rtol, 2.0 * P64_EPS * (double)(root.fe - root.fb));
  if (!isfinite(root.q) || fabs(root.q - t64) > rtol) return false;
  entropy = p64_entropy(e, n, kt, &root, &pub);
  if (!isfinite(entropy)) return false;
  for (int i = 0; i < n; ++i) occ[i] = p64_published(e, i, kt, &root, &pub);
  return true;
}

}  // namespace xtb
