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

int run_batch_lanes(const std::vector<Device*>& lanes, const HostBatch& hb, const RunOptions& opt,
                    HostOutputs& out, RunTiming* timing, std::string& error,
                    const std::vector<TopoCache*>* caches) {
  const int K = (int)std::min<int64_t>((int64_t)lanes.size(), hb.batch);
  if (K <= 1) {
    TopoCache* cache = caches ? (*caches)[0] : nullptr;
    if (cache) topo_cache_retag(cache, slice_tag(0, hb.batch, 1));
    return run_batch(lanes[0], hb, opt, out, timing, error, cache);
  }
  std::vector<int32_t> nao(hb.batch), ndim(hb.batch);
  if (batch_matrix_elements(hb, nullptr, error, nao.data(), ndim.data()) < 0) return 1;

  // cut points: equal shares of sum n^3, every slice non-empty
  std::vector<double> cost(hb.batch + 1, 0.0);
  for (int64_t s = 0; s < hb.batch; ++s) cost[s + 1] = cost[s] + (double)nao[s] * nao[s] * nao[s];
  std::vector<int64_t> cut(K + 1, 0);
  cut[K] = hb.batch;
  for (int k = 1; k < K; ++k) {
    const double target = cost[hb.batch] * k / K;
    int64_t c = std::lower_bound(cost.begin() + 1, cost.end(), target) - cost.begin();
    c = std::max<int64_t>(c, cut[k - 1] + 1);
    c = std::min<int64_t>(c, hb.batch - (K - k));
    cut[k] = c;
  }

  std::vector<int64_t> mat_prefix(hb.batch + 1, 0), state_prefix(hb.batch + 1, 0);
  for (int64_t s = 0; s < hb.batch; ++s) {
    mat_prefix[s + 1] = mat_prefix[s] + (int64_t)nao[s] * nao[s];
    state_prefix[s + 1] = state_prefix[s] + ndim[s];
  }
  std::vector<RunOptions> sopt(K, opt);

  std::vector<Slice> sl(K);
  for (int k = 0; k < K; ++k) {
    Slice& x = sl[k];
    x.s0 = cut[k];
    x.s1 = cut[k + 1];
    const int64_t a0 = hb.atom_offsets[x.s0], a1 = hb.atom_offsets[x.s1];
    x.atom_offsets.resize(x.s1 - x.s0 + 1);
    for (int64_t s = x.s0; s <= x.s1; ++s) x.atom_offsets[s - x.s0] = hb.atom_offsets[s] - a0;
    HostBatch& b = x.batch;
    b.batch = x.s1 - x.s0;
    b.atoms = a1 - a0;
    b.atom_offsets = x.atom_offsets.data();
    b.Z = hb.Z + a0;
    b.xyz = hb.xyz + 3 * a0;
    b.charges = hb.charges + x.s0;
    b.unpaired = hb.unpaired + x.s0;
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
    HostOutputs& o = x.out;
    const int64_t m0 = mat_prefix[x.s0];
    auto at = [](auto* p, int64_t off) { return p ? p + off : p; };
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
    o.scc_state = at(out.scc_state, state_prefix[x.s0]);
    sopt[k].warm_state = at(opt.warm_state, state_prefix[x.s0]);
    x.cache = caches ? (*caches)[k] : nullptr;
    if (x.cache) topo_cache_retag(x.cache, slice_tag(x.s0, x.s1, K));
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
