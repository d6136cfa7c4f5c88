// Host interface of the MI355X GFN2 executor (used by the C ABI layer).
#pragma once

#include <stddef.h>
#include <stdint.h>

#include <string>
#include <vector>

namespace xtb {

struct HostBatch {
  int64_t batch = 0;
  int64_t atoms = 0;
  int64_t points = 0;
  const int64_t* atom_offsets = nullptr;  // [B+1]
  const int32_t* Z = nullptr;             // [N]
  const double* xyz = nullptr;            // [3N]
  const double* charges = nullptr;        // [B]
  const int32_t* unpaired = nullptr;      // [B]
  const int64_t* point_offsets = nullptr; // [B+1] or null
  const double* pc_xyz = nullptr;         // [3P]
  const double* pc_q = nullptr;           // [P]
  const double* pc_gamma = nullptr;       // [P]
  const double* field = nullptr;          // [3B] uniform electric field per system, or null
  // external charge operator v = b + A q on the atomic charges: b [N] and
  // A (n_atoms^2 per system, row-major, packed in system order), each or null
  const double* pshift = nullptr;
  const double* presp = nullptr;
  const int32_t* spin = nullptr;          // [B] spin channels (1 or 2), or null (all 1)
};

struct RunOptions {
  int max_iterations = 250;
  double charge_tolerance = 1e-6;
  double energy_tolerance = 1e-8;
  double kt = 300.0 * 3.166808578545117e-6;
  int history = 8;
  double damping = 0.4;
  bool forces = true;
  // XTBLOOM_DETERMINISM_REPRODUCIBLE: order-independent (fixed-point) sums of
  // the scattered force contributions, so repeated calls agree bit for bit
  bool reproducible = false;
  // WARM start: the SCC multipole state to start from instead of the FRESH
  // SAD state, per system (shell charges, then 3 dipole and 6 quadrupole
  // components per atom), systems in batch order; null = FRESH
  const double* warm_state = nullptr;
  // size the device workspace for the batch and return before any kernel
  bool reserve_only = false;
  // batched path: the SCC rounds run as captured graphs of kGraphRounds
  // rounds each, driven by a device-resident round counter (the host reads
  // the active count once per graph, one graph ahead), and the mixed-precision
  // products; false: the plain FP64 path with one host check per round
  bool batched = false;
  // split-precision FP32 matrix-core density products: P in the SCC rounds
  // whose mixer residual is above 1e-3 (and the first round), W always
  bool split32 = false;
  // allocate nothing: the device workspace is the fixed per-GPU array of
  // kNoAllocWorkspaceBytes (a larger batch is rejected before any kernel),
  // host staging is pinned only when the context's pinned block is already
  // large enough, one lane, no per-launch events
  bool no_alloc = false;
};

// Largest AO count of one system; larger systems of a batch get per-system
// status NOT_SUPPORTED.  (Systems above 1024 AOs run kernel instantiations
// whose per-system work vectors live in global memory instead of LDS; a
// system of 8192 AOs holds about 20 n^2 doubles = 10.7 GB of device arena.)
constexpr int kMaxSystemAOs = 8192;

// Size of the per-GPU device workspace array behind RunOptions::no_alloc.
constexpr size_t kNoAllocWorkspaceBytes = size_t(64) << 20;

// Per-kernel-family timing record (layout of xtbloom_amd_kernel_stat_t).
struct KernelStat {
  int32_t launches;
  int32_t reserved;
  double total_ms;
  double flops;
  double bytes;
};
enum KernelKind {
  kKGeometry, kKIntegrals, kKCholesky, kKPotentials, kKFock, kKReduce, kKTridiag, kKFermi, kKBack,
  kKDensity, kKMulliken, kKEnergyMix, kKEnergyWeighted, kKForcePairs, kKForceAtoms, kKForceAtm,
  kKTridiagEigen, kKBacktransform, kKernelCount
};

// Every pointer is optional (host memory); failed systems receive NaN slices.
struct HostOutputs {
  double* energies = nullptr;       // [B]
  double* gradient = nullptr;       // [3N] dE/dR
  double* charges = nullptr;        // [N]
  double* pc_gradient = nullptr;    // [3P] dE/dR_point
  double* pc_dedq = nullptr;        // [P]  dE/dQ_point
  double* dipoles = nullptr;        // [3B]
  int32_t* iterations = nullptr;    // [B]
  uint8_t* converged = nullptr;     // [B]
  int32_t* status = nullptr;        // [B]
  // diagnostics (packed over the batch)
  double* dbg_overlap = nullptr;    // [sum n^2]
  double* dbg_h0 = nullptr;         // [sum n^2]
  double* dbg_density = nullptr;    // [sum n^2]
  double* dbg_cn = nullptr;         // [N]
  double* dbg_components = nullptr; // [B*8] core es2 es3 aes2 d4 pc rep atm
  double* dbg_levels = nullptr;     // [sum n] final-round orbital energies, ascending per system
  double* dbg_occ = nullptr;        // [sum n] their total occupations
  double* scc_state = nullptr;        // [sum (shells + 9 atoms)] final SCC multipole state
  KernelStat* kernel_stats = nullptr; // [kKernelCount]; enables per-launch events
  uint32_t kernel_stats_mask = 0;     // bit f: events for family f (0 = every family)
};

struct Device;
// Host topology of one batch (everything but the geometry), reusable by
// later calls with the same topology (fixed-topology plans).
struct TopoCache;
TopoCache* topo_cache_create();
void topo_cache_destroy(TopoCache* cache);
// Binds the cache to `tag` (which slice of which lane split); a new tag
// empties it.  Returns whether the cache holds a topology.
bool topo_cache_retag(TopoCache* cache, int64_t tag);

// Timing of the last run (milliseconds of device time per phase).
struct RunTiming {
  float setup_ms = 0.f, scc_ms = 0.f, force_ms = 0.f, total_ms = 0.f;
  int scc_rounds = 0;
  size_t workspace_bytes = 0;  // device workspace the call needs
};

Device* device_create(int device_id, void* stream, std::string& error);
void device_destroy(Device* dev);
int device_id(const Device* dev);
// returns an xtbloom status code; error receives a message on failure
// `cache` (optional) holds the host topology: when it is filled the call
// refreshes only the geometry in it, otherwise it builds and keeps it.  The
// caller guarantees that a filled cache belongs to the batch's topology.
int run_batch(Device* dev, const HostBatch& batch, const RunOptions& options, HostOutputs& out,
              RunTiming* timing, std::string& error, TopoCache* cache = nullptr);
// Packed-matrix element count and AO count limit checks used by validation;
// system_aos (optional, [B]) receives each system's AO count and
// system_state (optional, [B]) the length of its SCC multipole state (two
// channels of shell charges, dipoles and quadrupoles for an unrestricted
// system).
int64_t batch_matrix_elements(const HostBatch& batch, int32_t* max_ao, std::string& error,
                              int32_t* system_aos = nullptr, int32_t* system_state = nullptr);

// Runs a batch as up to `lanes.size()` contiguous, cost-balanced slices at
// the same time, one host thread and one device stream per slice (lanes[0]
// runs in the calling thread).  Results are identical to run_batch on the
// whole batch: every system is computed independently.  Returns an xtbloom
// status code like run_batch; warm_state and scc_state are cut per slice.
// caches (optional) holds one topology cache per lane; a cache is reused only
// for the same slice of the same lane count.
// Systems above kMaxSystemAOs are left out of the slices and reported with
// per-system status NOT_SUPPORTED and NaN slices.
int run_batch_lanes(const std::vector<Device*>& lanes, const HostBatch& batch,
                    const RunOptions& options, HostOutputs& out, RunTiming* timing,
                    std::string& error, const std::vector<TopoCache*>* caches = nullptr);
// Lane count for a batch: min(configured lanes, batch / kLaneMinSystems), at
// least 1.  The configured count is XTBLOOM_AMD_LANES (default 1).
int lanes_for_batch(int64_t batch);

}  // namespace xtb
