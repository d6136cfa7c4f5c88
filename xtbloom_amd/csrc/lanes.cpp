// Concurrent execution lanes for one batch.
//
// Most kernels of the GFN2 path are bound by the largest system of the batch
// (one workgroup per system for the Householder reduction, the eigensolver and
// the SCC mixer) while the streaming kernels (Fock build, Mulliken sums,
// GEMMs) scale with the whole batch.  Cutting the batch into a few contiguous
// slices with about equal sum of n^3 and running each slice's complete
// pipeline on its own stream lets the streaming work of one slice fill the
// compute units that the latency-bound kernels of another slice leave idle,
// and overlaps one slice's host bookkeeping with another slice's kernels.
// On the 512-system ragged alkane batch two lanes run 3-5 % faster than one.
// Lanes are opt-in (XTBLOOM_AMD_LANES): every lane holds a stream, and once a
// process has more busy streams than the runtime's hardware queues (four by
// default) streams share queues and serialize, which costs far more than the
// lanes gain; one lane is the default.
// Every system is computed independently of the others, so the results equal
// those of one run over the whole batch, bit for bit apart from the order of
// the floating-point atomic additions of the gradient kernels.
#include <stdlib.h>

#include <algorithm>
#include <limits>
#include <new>
#include <thread>
#include <vector>

#include "executor.hpp"

namespace xtb {

namespace {

constexpr int kDefaultLanes = 1;
constexpr int kMaxLanes = 4;
constexpr int64_t kLaneMinSystems = 32;

struct Slice {
  int64_t s0, s1;  // systems [s0, s1)
  std::vector<int64_t> atom_offsets, point_offsets;
  HostBatch batch;
  HostOutputs out;
  std::vector<KernelStat> stats;
  RunTiming timing;
  std::string error;
  TopoCache* cache = nullptr;
  int rc = 0;
};

// exact identity of a slice: s0, s1 < 2^29 (batch sizes are below 2^28), K <= 4
int64_t slice_tag(int64_t s0, int64_t s1, int K) { return (s0 << 34) | (s1 << 4) | K; }


int run_guarded(Device* dev, Slice& sl, const RunOptions& opt) {
  try {
    return run_batch(dev, sl.batch, opt, sl.out, &sl.timing, sl.error, sl.cache);
  } catch (const std::bad_alloc&) {
    sl.error = "host allocation failed";
    return 4;  // XTBLOOM_STATUS_ALLOCATION_FAILED
  } catch (...) {
    sl.error = "unexpected exception in the ROCm executor";
    return 6;  // XTBLOOM_STATUS_INTERNAL_ERROR
  }
}

}  // namespace

int lanes_for_batch(int64_t batch) {
  int want = kDefaultLanes;
  if (const char* env = getenv("XTBLOOM_AMD_LANES")) {
    const int v = atoi(env);
    if (v >= 1) want = v;
  }
  want = std::min(want, kMaxLanes);
  const int64_t fit = batch / kLaneMinSystems;
  return (int)std::max<int64_t>(1, std::min<int64_t>(want, fit));
}

namespace {

// Systems of each lane, ascending batch order inside a lane.
using LaneMembers = std::vector<std::vector<int64_t>>;

// K contiguous runs of `kept` with about equal sum of n^3, every run non-empty.
LaneMembers contiguous_split(const std::vector<int32_t>& nao, const std::vector<int64_t>& kept, int K) {
  const int64_t B = (int64_t)kept.size();
  std::vector<double> cost(B + 1, 0.0);
  for (int64_t i = 0; i < B; ++i) {
    const double n = nao[kept[i]];
    cost[i + 1] = cost[i] + n * n * n;
  }
  std::vector<int64_t> cut(K + 1, 0);
  cut[K] = B;
  for (int k = 1; k < K; ++k) {
    const double target = cost[B] * k / K;
    int64_t c = std::lower_bound(cost.begin() + 1, cost.end(), target) - cost.begin();
    c = std::max<int64_t>(c, cut[k - 1] + 1);
    c = std::min<int64_t>(c, B - (K - k));
    cut[k] = c;
  }
  LaneMembers m(K);
  for (int k = 0; k < K; ++k)
    for (int64_t i = cut[k]; i < cut[k + 1]; ++i) m[k].push_back(kept[i]);
  return m;
}

bool is_contiguous(const std::vector<int64_t>& m) {
  for (size_t i = 1; i < m.size(); ++i)
    if (m[i] != m[i - 1] + 1) return false;
  return true;
}

// Slice-owned copies of the inputs and outputs of a non-contiguous lane.
struct Gathered {
  std::vector<int64_t> atom_offsets, point_offsets;
  std::vector<int32_t> Z, unpaired, spin, iterations, status;
  std::vector<double> xyz, charges, pc_xyz, pc_q, pc_gamma, field, pshift, presp, warm;
  std::vector<double> energies, gradient, atom_charges, pc_gradient, pc_dedq, dipoles, scc_state,
      dbg_overlap, dbg_h0, dbg_density, dbg_cn, dbg_components, dbg_levels, dbg_occ;
  std::vector<uint8_t> converged;
};

// Outputs of the systems above the AO limit: NaN slices, status
// NOT_SUPPORTED, no iterations (the other outputs of a call are untouched
// by this).
void fill_unsupported(const HostBatch& hb, const std::vector<int32_t>& nao,
                      const std::vector<int32_t>& ndim, HostOutputs& out, int limit) {
  const double nan = std::numeric_limits<double>::quiet_NaN();
  int64_t state = 0, ao = 0;
  for (int64_t s = 0; s < hb.batch; ++s) {
    const int64_t ns = ndim[s];
    if (nao[s] > limit) {
      const int64_t a0 = hb.atom_offsets[s], a1 = hb.atom_offsets[s + 1];
      if (out.status) out.status[s] = 3;  // XTBLOOM_STATUS_NOT_SUPPORTED
      if (out.iterations) out.iterations[s] = 0;
      if (out.converged) out.converged[s] = 0;
      if (out.energies) out.energies[s] = nan;
      for (int64_t a = a0; a < a1; ++a) {
        if (out.charges) out.charges[a] = nan;
        if (out.gradient)
          for (int c = 0; c < 3; ++c) out.gradient[3 * a + c] = nan;
        if (out.dbg_cn) out.dbg_cn[a] = nan;
      }
      if (out.dipoles)
        for (int c = 0; c < 3; ++c) out.dipoles[3 * s + c] = nan;
      if (hb.point_offsets)
        for (int64_t p = hb.point_offsets[s]; p < hb.point_offsets[s + 1]; ++p) {
          if (out.pc_gradient)
            for (int c = 0; c < 3; ++c) out.pc_gradient[3 * p + c] = nan;
          if (out.pc_dedq) out.pc_dedq[p] = nan;
        }
      if (out.dbg_components)
        for (int c = 0; c < 8; ++c) out.dbg_components[8 * s + c] = nan;
      for (int64_t i = ao; i < ao + nao[s]; ++i) {
        if (out.dbg_levels) out.dbg_levels[i] = nan;
        if (out.dbg_occ) out.dbg_occ[i] = nan;
      }
      if (out.scc_state)
        for (int64_t i = 0; i < ns; ++i) out.scc_state[state + i] = nan;
    }
    state += ns;
    ao += nao[s];
  }
}

}  // namespace

int run_batch_lanes(const std::vector<Device*>& lanes, const HostBatch& hb, const RunOptions& opt,
                    HostOutputs& out, RunTiming* timing, std::string& error,
                    const std::vector<TopoCache*>* caches) {
  std::vector<int32_t> nao(hb.batch), ndim(hb.batch);
  int32_t max_ao = 0;
  if (batch_matrix_elements(hb, &max_ao, error, nao.data(), ndim.data()) < 0) return 1;
  // Systems above the eigensolver's AO limit are not run: they get
  // per-system status NOT_SUPPORTED with NaN slices, the rest of the batch is
  // computed as usual.
  std::vector<int64_t> kept;
  kept.reserve(hb.batch);
  for (int64_t s = 0; s < hb.batch; ++s)
    if (nao[s] <= kMaxSystemAOs) kept.push_back(s);
  const bool excluded = (int64_t)kept.size() < hb.batch;
  const int K = (int)std::min<int64_t>((int64_t)lanes.size(), (int64_t)kept.size());
  if (K <= 1 && !excluded) {
    TopoCache* cache = caches ? (*caches)[0] : nullptr;
    if (cache) topo_cache_retag(cache, slice_tag(0, hb.batch, 1));
    return run_batch(lanes[0], hb, opt, out, timing, error, cache);
  }
  if (excluded) fill_unsupported(hb, nao, ndim, out, kMaxSystemAOs);
  if (K == 0) {
    if (timing) *timing = RunTiming();
    if (out.kernel_stats)
      for (int f = 0; f < kKernelCount; ++f) out.kernel_stats[f] = KernelStat{0, 0, 0.0, 0.0, 0.0};
    return 0;
  }
  const LaneMembers members = contiguous_split(nao, kept, K);

  std::vector<int64_t> mat_prefix(hb.batch + 1, 0), state_prefix(hb.batch + 1, 0), ao_prefix(hb.batch + 1, 0),
      resp_prefix(hb.batch + 1, 0);
  for (int64_t s = 0; s < hb.batch; ++s) {
    const int64_t na = hb.atom_offsets[s + 1] - hb.atom_offsets[s];
    resp_prefix[s + 1] = resp_prefix[s] + na * na;
    mat_prefix[s + 1] = mat_prefix[s] + (int64_t)nao[s] * nao[s];
    state_prefix[s + 1] = state_prefix[s] + ndim[s];
    ao_prefix[s + 1] = ao_prefix[s] + nao[s];
  }
  std::vector<RunOptions> sopt(K, opt);
  std::vector<Slice> sl(K);
  std::vector<Gathered> gat(K);
  auto at = [](auto* p, int64_t off) { return p ? p + off : p; };
  for (int k = 0; k < K; ++k) {
    Slice& x = sl[k];
    const std::vector<int64_t>& m = members[k];
    x.s0 = m.front();
    x.s1 = m.back() + 1;
    HostBatch& b = x.batch;
    HostOutputs& o = x.out;
    b.batch = (int64_t)m.size();
    if (is_contiguous(m)) {
      const int64_t a0 = hb.atom_offsets[x.s0], a1 = hb.atom_offsets[x.s1];
      x.atom_offsets.resize(x.s1 - x.s0 + 1);
      for (int64_t s = x.s0; s <= x.s1; ++s) x.atom_offsets[s - x.s0] = hb.atom_offsets[s] - a0;
      b.atoms = a1 - a0;
      b.atom_offsets = x.atom_offsets.data();
      b.Z = hb.Z + a0;
      b.xyz = hb.xyz + 3 * a0;
      b.charges = hb.charges + x.s0;
      b.unpaired = hb.unpaired + x.s0;
      b.field = at(hb.field, 3 * x.s0);
      b.spin = at(hb.spin, x.s0);
      b.pshift = at(hb.pshift, a0);
      b.presp = at(hb.presp, resp_prefix[x.s0]);
      int64_t p0 = 0;
      if (hb.point_offsets) {
        p0 = hb.point_offsets[x.s0];
        const int64_t p1 = hb.point_offsets[x.s1];
        x.point_offsets.resize(x.s1 - x.s0 + 1);
        for (int64_t s = x.s0; s <= x.s1; ++s) x.point_offsets[s - x.s0] = hb.point_offsets[s] - p0;
        b.points = p1 - p0;
        b.point_offsets = x.point_offsets.data();
        b.pc_xyz = hb.pc_xyz + 3 * p0;
        b.pc_q = hb.pc_q + p0;
        b.pc_gamma = hb.pc_gamma + p0;
      }
      const int64_t m0 = mat_prefix[x.s0];
      o.energies = at(out.energies, x.s0);
      o.gradient = at(out.gradient, 3 * a0);
      o.charges = at(out.charges, a0);
      o.pc_gradient = at(out.pc_gradient, 3 * p0);
      o.pc_dedq = at(out.pc_dedq, p0);
      o.dipoles = at(out.dipoles, 3 * x.s0);
      o.iterations = at(out.iterations, x.s0);
      o.converged = at(out.converged, x.s0);
      o.status = at(out.status, x.s0);
      o.dbg_overlap = at(out.dbg_overlap, m0);
      o.dbg_h0 = at(out.dbg_h0, m0);
      o.dbg_density = at(out.dbg_density, m0);
      o.dbg_cn = at(out.dbg_cn, a0);
      o.dbg_components = at(out.dbg_components, 8 * x.s0);
      o.dbg_levels = at(out.dbg_levels, ao_prefix[x.s0]);
      o.dbg_occ = at(out.dbg_occ, ao_prefix[x.s0]);
      o.scc_state = at(out.scc_state, state_prefix[x.s0]);
      sopt[k].warm_state = at(opt.warm_state, state_prefix[x.s0]);
    } else {
      // gather the members' inputs; outputs land in slice buffers and are
      // scattered back once every lane has finished
      Gathered& g = gat[k];
      g.atom_offsets.assign(1, 0);
      if (hb.point_offsets) g.point_offsets.assign(1, 0);
      int64_t nmat = 0, nstate = 0, nlev = 0;
      for (int64_t s : m) {
        const int64_t a0 = hb.atom_offsets[s], a1 = hb.atom_offsets[s + 1];
        g.atom_offsets.push_back(g.atom_offsets.back() + (a1 - a0));
        g.Z.insert(g.Z.end(), hb.Z + a0, hb.Z + a1);
        g.xyz.insert(g.xyz.end(), hb.xyz + 3 * a0, hb.xyz + 3 * a1);
        g.charges.push_back(hb.charges[s]);
        g.unpaired.push_back(hb.unpaired[s]);
        if (hb.field) g.field.insert(g.field.end(), hb.field + 3 * s, hb.field + 3 * s + 3);
        if (hb.spin) g.spin.push_back(hb.spin[s]);
        if (hb.pshift) g.pshift.insert(g.pshift.end(), hb.pshift + a0, hb.pshift + a1);
        if (hb.presp) g.presp.insert(g.presp.end(), hb.presp + resp_prefix[s], hb.presp + resp_prefix[s + 1]);
        if (hb.point_offsets) {
          const int64_t p0 = hb.point_offsets[s], p1 = hb.point_offsets[s + 1];
          g.point_offsets.push_back(g.point_offsets.back() + (p1 - p0));
          g.pc_xyz.insert(g.pc_xyz.end(), hb.pc_xyz + 3 * p0, hb.pc_xyz + 3 * p1);
          g.pc_q.insert(g.pc_q.end(), hb.pc_q + p0, hb.pc_q + p1);
          g.pc_gamma.insert(g.pc_gamma.end(), hb.pc_gamma + p0, hb.pc_gamma + p1);
        }
        if (opt.warm_state)
          g.warm.insert(g.warm.end(), opt.warm_state + state_prefix[s], opt.warm_state + state_prefix[s + 1]);
        nmat += (int64_t)nao[s] * nao[s];
        nstate += ndim[s];
        nlev += nao[s];
      }
      const int64_t nb = b.batch, na = g.atom_offsets.back();
      const int64_t np = hb.point_offsets ? g.point_offsets.back() : 0;
      b.atoms = na;
      b.atom_offsets = g.atom_offsets.data();
      b.Z = g.Z.data();
      b.xyz = g.xyz.data();
      b.charges = g.charges.data();
      b.unpaired = g.unpaired.data();
      b.field = hb.field ? g.field.data() : nullptr;
      b.spin = hb.spin ? g.spin.data() : nullptr;
      b.pshift = hb.pshift ? g.pshift.data() : nullptr;
      b.presp = hb.presp ? g.presp.data() : nullptr;
      if (hb.point_offsets) {
        b.points = np;
        b.point_offsets = g.point_offsets.data();
        b.pc_xyz = g.pc_xyz.data();
        b.pc_q = g.pc_q.data();
        b.pc_gamma = g.pc_gamma.data();
      }
      auto own = [](auto& vec, const void* want, int64_t n) {
        if (want == nullptr) return (decltype(vec.data()))nullptr;
        vec.assign((size_t)n, {});
        return vec.data();
      };
      o.energies = own(g.energies, out.energies, nb);
      o.gradient = own(g.gradient, out.gradient, 3 * na);
      o.charges = own(g.atom_charges, out.charges, na);
      o.pc_gradient = own(g.pc_gradient, out.pc_gradient, 3 * np);
      o.pc_dedq = own(g.pc_dedq, out.pc_dedq, np);
      o.dipoles = own(g.dipoles, out.dipoles, 3 * nb);
      o.iterations = own(g.iterations, out.iterations, nb);
      o.converged = own(g.converged, out.converged, nb);
      o.status = own(g.status, out.status, nb);
      o.dbg_overlap = own(g.dbg_overlap, out.dbg_overlap, nmat);
      o.dbg_h0 = own(g.dbg_h0, out.dbg_h0, nmat);
      o.dbg_density = own(g.dbg_density, out.dbg_density, nmat);
      o.dbg_cn = own(g.dbg_cn, out.dbg_cn, na);
      o.dbg_components = own(g.dbg_components, out.dbg_components, 8 * nb);
      o.dbg_levels = own(g.dbg_levels, out.dbg_levels, nlev);
      o.dbg_occ = own(g.dbg_occ, out.dbg_occ, nlev);
      o.scc_state = own(g.scc_state, out.scc_state, nstate);
      sopt[k].warm_state = opt.warm_state ? g.warm.data() : nullptr;
    }
    x.cache = caches ? (*caches)[k] : nullptr;
    // a slice with systems left out is a gathered slice: its tag differs
    // from the contiguous slice over the same range
    if (x.cache) topo_cache_retag(x.cache, is_contiguous(m) ? slice_tag(x.s0, x.s1, K) : -slice_tag(x.s0, x.s1, K) - 1);
    if (out.kernel_stats) {
      x.stats.assign(kKernelCount, KernelStat{0, 0, 0.0, 0.0, 0.0});
      o.kernel_stats = x.stats.data();
      o.kernel_stats_mask = out.kernel_stats_mask;
    }
  }

  std::vector<std::thread> workers;
  workers.reserve(K - 1);
  int spawned = 1;
  try {
    for (int k = 1; k < K; ++k) {
      workers.emplace_back([&, k]() { sl[k].rc = run_guarded(lanes[k], sl[k], sopt[k]); });
      ++spawned;
    }
  } catch (...) {
    // could not start a thread: the remaining slices run in this thread
  }
  sl[0].rc = run_guarded(lanes[0], sl[0], sopt[0]);
  for (int k = spawned; k < K; ++k) sl[k].rc = run_guarded(lanes[k], sl[k], sopt[k]);
  for (auto& w : workers) w.join();

  for (int k = 0; k < K; ++k) {
    if (sl[k].rc != 0) {
      error = sl[k].error;
      return sl[k].rc;
    }
  }
  // scatter the gathered lanes' outputs to their systems
  for (int k = 0; k < K; ++k) {
    const std::vector<int64_t>& m = members[k];
    if (is_contiguous(m)) continue;
    const Gathered& g = gat[k];
    int64_t ga = 0, gp = 0, gm = 0, gs = 0, gl = 0;
    for (size_t i = 0; i < m.size(); ++i) {
      const int64_t s = m[i];
      const int64_t a0 = hb.atom_offsets[s], na = hb.atom_offsets[s + 1] - a0;
      const int64_t p0 = hb.point_offsets ? hb.point_offsets[s] : 0;
      const int64_t np = hb.point_offsets ? hb.point_offsets[s + 1] - p0 : 0;
      const int64_t nm = (int64_t)nao[s] * nao[s], ns = ndim[s];
      auto put = [](auto* dst, const auto& src, int64_t from, int64_t n) {
        if (dst) std::copy(src.begin() + from, src.begin() + from + n, dst);
      };
      put(at(out.energies, s), g.energies, (int64_t)i, 1);
      put(at(out.gradient, 3 * a0), g.gradient, 3 * ga, 3 * na);
      put(at(out.charges, a0), g.atom_charges, ga, na);
      put(at(out.pc_gradient, 3 * p0), g.pc_gradient, 3 * gp, 3 * np);
      put(at(out.pc_dedq, p0), g.pc_dedq, gp, np);
      put(at(out.dipoles, 3 * s), g.dipoles, 3 * (int64_t)i, 3);
      put(at(out.iterations, s), g.iterations, (int64_t)i, 1);
      put(at(out.converged, s), g.converged, (int64_t)i, 1);
      put(at(out.status, s), g.status, (int64_t)i, 1);
      put(at(out.dbg_overlap, mat_prefix[s]), g.dbg_overlap, gm, nm);
      put(at(out.dbg_h0, mat_prefix[s]), g.dbg_h0, gm, nm);
      put(at(out.dbg_density, mat_prefix[s]), g.dbg_density, gm, nm);
      put(at(out.dbg_cn, a0), g.dbg_cn, ga, na);
      put(at(out.dbg_components, 8 * s), g.dbg_components, 8 * (int64_t)i, 8);
      put(at(out.scc_state, state_prefix[s]), g.scc_state, gs, ns);
      put(at(out.dbg_levels, ao_prefix[s]), g.dbg_levels, gl, nao[s]);
      put(at(out.dbg_occ, ao_prefix[s]), g.dbg_occ, gl, nao[s]);
      gl += nao[s];
      ga += na;
      gp += np;
      gm += nm;
      gs += ns;
    }
  }
  if (out.kernel_stats) {
    for (int f = 0; f < kKernelCount; ++f) {
      KernelStat sum{0, 0, 0.0, 0.0, 0.0};
      for (const Slice& x : sl) {
        sum.launches += x.stats[f].launches;
        sum.total_ms += x.stats[f].total_ms;
        sum.flops += x.stats[f].flops;
        sum.bytes += x.stats[f].bytes;
      }
      out.kernel_stats[f] = sum;
    }
  }
  if (timing) {
    RunTiming t;
    for (const Slice& x : sl) {
      t.setup_ms = std::max(t.setup_ms, x.timing.setup_ms);
      t.scc_ms = std::max(t.scc_ms, x.timing.scc_ms);
      t.force_ms = std::max(t.force_ms, x.timing.force_ms);
      t.total_ms = std::max(t.total_ms, x.timing.total_ms);
      t.scc_rounds = std::max(t.scc_rounds, x.timing.scc_rounds);
      t.workspace_bytes += x.timing.workspace_bytes;
    }
    *timing = t;
  }
  return 0;
}

}  // namespace xtb
