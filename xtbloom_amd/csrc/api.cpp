// xtbloom C ABI served by the MI355X GFN2 executor.
//
// Entry-point semantics follow the reference ABI (src/api.cpp, include/xtbloom/
// xtbloom.h): versioned struct prefixes, thread-local last error, validation of
// every descriptor before any output byte is written, per-system numerical
// failures reported as data (SUCCESS + NaN slices + per_system_status), and no
// exception crosses the boundary.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <memory>
#include <mutex>
#include <new>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/xtbloom/xtbloom.h"
#include "../../include/xtbloom_amd/extensions.h"
#include "executor.hpp"

static_assert(sizeof(xtb::KernelStat) == sizeof(xtbloom_amd_kernel_stat_t), "kernel stat layout");
static_assert(offsetof(xtb::KernelStat, bytes) == offsetof(xtbloom_amd_kernel_stat_t, bytes),
              "kernel stat layout");
static_assert((int)xtb::kKernelCount == (int)XTBLOOM_AMD_KERNEL_COUNT, "kernel family count");
static_assert(xtb::kNoAllocWorkspaceBytes == XTBLOOM_AMD_NO_ALLOC_WORKSPACE_BYTES, "no-allocation workspace size");

// Identity of a call for strict WARM starts and fixed-topology plans
// (reference xtbloom.h:502-520): the topology (atom offsets, element numbers,
// point-charge offsets and hardnesses, spin channels), molecular charges,
// unpaired electrons and the compute policy.  Geometry (atom and point-charge
// positions, point-charge values) is not part of it.
struct CallIdentity {
  std::vector<char> atom_offsets, Z, charges, unpaired, point_offsets, pc_gamma, spin;
  uint32_t flags = 0;
  int32_t model = 0, max_iterations = 0, history = 0, determinism = 0;
  double charge_tolerance = 0.0, energy_tolerance = 0.0, kt = 0.0, damping = 0.0;
  bool operator==(const CallIdentity& o) const {
    // doubles compare by their bits: the identity is exact
    auto same = [](double a, double b) { return std::memcmp(&a, &b, sizeof(double)) == 0; };
    return atom_offsets == o.atom_offsets && Z == o.Z && charges == o.charges && unpaired == o.unpaired &&
           point_offsets == o.point_offsets && pc_gamma == o.pc_gamma && spin == o.spin && flags == o.flags &&
           model == o.model && max_iterations == o.max_iterations && history == o.history &&
           determinism == o.determinism && same(charge_tolerance, o.charge_tolerance) &&
           same(energy_tolerance, o.energy_tolerance) && same(kt, o.kt) && same(damping, o.damping);
  }
};

// Converged SCC multipole state of the latest fully converged call, the
// starting point of the next strict WARM call with the same identity
// (reference gfn2_cpu_execution.cpp:829-838, 1313-1325, 1470-1487).
struct WarmCheckpoint {
  bool valid = false;
  CallIdentity id;
  std::vector<double> state;
};

struct xtbloom_context {
  xtb::Device* device = nullptr;
  // additional execution lanes (own streams and arenas) for large batches,
  // created on first use
  std::vector<xtb::Device*> lanes;
  WarmCheckpoint warm;  // checkpoint of xtbloom_compute calls
  // host topology of the previous xtbloom_compute call, reused when the next
  // call has the same topology (only its geometry is then refreshed)
  std::vector<xtb::TopoCache*> topo;
  CallIdentity topo_id;
  ~xtbloom_context() {
    for (xtb::TopoCache* c : topo) xtb::topo_cache_destroy(c);
  }
  std::mutex lock;  // one context is one serialized transaction domain (reference backend.hpp:28-32)
};

// The inputs the packed host topology depends on: atom offsets, element
// numbers, molecular charges, unpaired electrons, point-charge offsets and
// gammas.
// Asynchronous submission owner (reference xtbloom.h:41-47, 836-878).  This
// backend settles every accepted submission inline: enqueue runs the call and
// leaves the request COMPLETE, which the reference ABI allows ("or COMPLETE
// when the backend must settle it inline"); admission errors return directly
// and leave the request unchanged.
struct xtbloom_request {
  xtbloom_context_t* context = nullptr;
  int32_t state = XTBLOOM_REQUEST_IDLE;
  int32_t completion_status = XTBLOOM_STATUS_SUCCESS;
  uint32_t result_flags = 0;
  std::string error;
};

bool same_topology(const CallIdentity& a, const CallIdentity& b) {
  return a.atom_offsets == b.atom_offsets && a.Z == b.Z && a.charges == b.charges &&
         a.unpaired == b.unpaired && a.point_offsets == b.point_offsets && a.pc_gamma == b.pc_gamma;
}

// Fixed-topology plan (reference xtbloom.h:939-991): the identity it was
// created for, its own WARM checkpoint, the host topology of each lane slice
// and the workspace sizes reserved at creation.
struct xtbloom_plan {
  xtbloom_context_t* context = nullptr;
  CallIdentity id;
  WarmCheckpoint warm;
  std::vector<xtb::TopoCache*> caches;
  uint64_t device_bytes = 0, host_bytes = 0;
  ~xtbloom_plan() {
    for (xtb::TopoCache* c : caches) xtb::topo_cache_destroy(c);
  }
};

namespace {

thread_local std::string g_last_error;

xtbloom_status_t fail(xtbloom_status_t status, std::string message) {
  g_last_error = std::move(message);
  return status;
}

template <typename T>
xtbloom_status_t init_struct(T* value, size_t caller_size, size_t minimum, const char* name) {
  if (value == nullptr) return fail(XTBLOOM_STATUS_INVALID_ARGUMENT, std::string(name) + " is NULL");
  if (caller_size < minimum || caller_size > std::numeric_limits<uint32_t>::max())
    return fail(XTBLOOM_STATUS_INVALID_ARGUMENT,
                std::string(name) + " size is smaller than ABI v1 or exceeds uint32_t");
  std::memset(value, 0, std::min(caller_size, sizeof(T)));
  value->struct_size = static_cast<uint32_t>(caller_size);
  value->api_version = XTBLOOM_API_VERSION;
  g_last_error.clear();
  return XTBLOOM_STATUS_SUCCESS;
}

template <typename T>
bool valid_header(const T* v, size_t minimum) {
  return v != nullptr && v->struct_size >= minimum && v->api_version == XTBLOOM_API_VERSION;
}

bool buffer_active(const xtbloom_const_buffer_t& b) { return b.data != nullptr; }

// Copies a caller input buffer (host or ROCm device) into host memory.
xtbloom_status_t stage_input(const xtbloom_const_buffer_t& b, size_t need, const char* name,
                             std::vector<char>& dst, bool required) {
  if (b.data == nullptr) {
    if (required && need > 0) return fail(XTBLOOM_STATUS_INVALID_ARGUMENT, std::string(name) + " is NULL");
    dst.clear();
    return XTBLOOM_STATUS_SUCCESS;
  }
  if (b.size_bytes < need)
    return fail(XTBLOOM_STATUS_INVALID_ARGUMENT, std::string(name) + " buffer is too small");
  if (b.reserved != 0) return fail(XTBLOOM_STATUS_INVALID_ARGUMENT, std::string(name) + " reserved field must be zero");
  dst.resize(need);
  if (need == 0) return XTBLOOM_STATUS_SUCCESS;
  if (b.memory_space == XTBLOOM_MEMORY_HOST) {
    std::memcpy(dst.data(), b.data, need);
  } else if (b.memory_space == XTBLOOM_MEMORY_ROCM_DEVICE) {
    if (hipMemcpy(dst.data(), b.data, need, hipMemcpyDeviceToHost) != hipSuccess)
      return fail(XTBLOOM_STATUS_INVALID_ARGUMENT, std::string(name) + " is not a readable ROCm device buffer");
  } else {
    return fail(XTBLOOM_STATUS_INVALID_ARGUMENT,
                std::string(name) + " memory space is not HOST or ROCM_DEVICE for this backend");
  }
  return XTBLOOM_STATUS_SUCCESS;
}

xtbloom_status_t check_output(const xtbloom_buffer_t& b, size_t need, const char* name, bool required) {
  if (b.data == nullptr) {
    if (required) return fail(XTBLOOM_STATUS_INVALID_ARGUMENT, std::string(name) + " output is required");
    return XTBLOOM_STATUS_SUCCESS;
  }
  if (b.size_bytes < need) return fail(XTBLOOM_STATUS_INVALID_ARGUMENT, std::string(name) + " output is too small");
  if (b.reserved != 0) return fail(XTBLOOM_STATUS_INVALID_ARGUMENT, std::string(name) + " reserved field must be zero");
  if (b.memory_space != XTBLOOM_MEMORY_HOST && b.memory_space != XTBLOOM_MEMORY_ROCM_DEVICE)
    return fail(XTBLOOM_STATUS_INVALID_ARGUMENT, std::string(name) + " memory space is not HOST or ROCM_DEVICE");
  return XTBLOOM_STATUS_SUCCESS;
}

xtbloom_status_t publish(const xtbloom_buffer_t& b, const void* src, size_t bytes) {
  if (b.data == nullptr || bytes == 0) return XTBLOOM_STATUS_SUCCESS;
  if (b.memory_space == XTBLOOM_MEMORY_HOST) {
    std::memcpy(b.data, src, bytes);
  } else if (hipMemcpy(b.data, src, bytes, hipMemcpyHostToDevice) != hipSuccess) {
    return fail(XTBLOOM_STATUS_INTERNAL_ERROR, "failed to publish into a ROCm device output buffer");
  }
  return XTBLOOM_STATUS_SUCCESS;
}

struct Staged {
  std::vector<char> atom_offsets, Z, xyz, charges, unpaired, poff, pxyz, pq, pg, spin;
};

// result may be NULL (plan creation: no outputs are checked).
xtbloom_status_t validate_and_stage(const xtbloom_batch_t* batch, const xtbloom_compute_options_t* options,
                                    const xtbloom_batch_result_t* result, Staged& st,
                                    xtb::HostBatch& hb, xtb::RunOptions& ro, uint32_t& flags, bool& warm) {
  warm = false;
  if (!valid_header(batch, XTBLOOM_BATCH_V1_SIZE))
    return fail(XTBLOOM_STATUS_INVALID_ARGUMENT, "batch is NULL, too small, or has a different api_version");
  if (!valid_header(options, XTBLOOM_COMPUTE_OPTIONS_V1_SIZE))
    return fail(XTBLOOM_STATUS_INVALID_ARGUMENT, "compute options are NULL, too small, or have a different api_version");
  if (result != nullptr && !valid_header(result, XTBLOOM_BATCH_RESULT_V1_SIZE))
    return fail(XTBLOOM_STATUS_INVALID_ARGUMENT, "batch result is NULL, too small, or has a different api_version");
  const int64_t B = batch->batch_size, N = batch->total_atoms, P = batch->total_point_charges;
  if (B <= 0 || N <= 0 || N < B || P < 0 || B > (int64_t)1 << 28 || N > (int64_t)1 << 30 || P > (int64_t)1 << 30)
    return fail(XTBLOOM_STATUS_INVALID_ARGUMENT, "batch_size, total_atoms or total_point_charges is out of range");
  // options
  if (options->model == XTBLOOM_MODEL_GFN1_XTB)
    return fail(XTBLOOM_STATUS_NOT_SUPPORTED, "the ROCm backend implements GFN2-xTB only");
  if (options->model != XTBLOOM_MODEL_GFN2_XTB)
    return fail(XTBLOOM_STATUS_INVALID_ARGUMENT, "unknown model");
  flags = options->flags;
  const uint32_t known = XTBLOOM_COMPUTE_ENERGY | XTBLOOM_COMPUTE_FORCES | XTBLOOM_COMPUTE_ATOMIC_CHARGES |
                         XTBLOOM_COMPUTE_POINT_CHARGE_FORCES | XTBLOOM_COMPUTE_DIPOLE_MOMENTS;
  if ((flags & ~known) != 0) return fail(XTBLOOM_STATUS_INVALID_ARGUMENT, "unknown compute flags");
  if (options->max_scc_iterations <= 0) return fail(XTBLOOM_STATUS_INVALID_ARGUMENT, "max_scc_iterations must be positive");
  if (!(options->charge_tolerance > 0.0) || !std::isfinite(options->charge_tolerance) ||
      !(options->energy_tolerance > 0.0) || !std::isfinite(options->energy_tolerance))
    return fail(XTBLOOM_STATUS_INVALID_ARGUMENT, "SCC tolerances must be finite and positive");
  if (!(options->electronic_temperature >= 0.0) || !std::isfinite(options->electronic_temperature))
    return fail(XTBLOOM_STATUS_INVALID_ARGUMENT, "electronic_temperature must be finite and nonnegative");
  // reserved fields and enum ranges (reference validation.cpp:426-461)
  if (options->reserved != 0)
    return fail(XTBLOOM_STATUS_INVALID_ARGUMENT, "compute options reserved field must be zero");
  ro.max_iterations = options->max_scc_iterations;
  ro.charge_tolerance = options->charge_tolerance;
  ro.energy_tolerance = options->energy_tolerance;
  ro.kt = options->electronic_temperature;
  if (options->struct_size >= XTBLOOM_COMPUTE_OPTIONS_V2_SIZE) {
    if (options->scc_start_mode != XTBLOOM_SCC_START_FRESH && options->scc_start_mode != XTBLOOM_SCC_START_WARM)
      return fail(XTBLOOM_STATUS_INVALID_ARGUMENT, "scc_start_mode must be XTBLOOM_SCC_START_FRESH or XTBLOOM_SCC_START_WARM");
    if (options->reserved_v2 != 0)
      return fail(XTBLOOM_STATUS_INVALID_ARGUMENT, "compute options reserved_v2 field must be zero");
    warm = options->scc_start_mode == XTBLOOM_SCC_START_WARM;
  }
  if (options->struct_size >= XTBLOOM_COMPUTE_OPTIONS_V3_SIZE) {
    if (options->scc_mixer != XTBLOOM_SCC_MIXER_MODIFIED_BROYDEN)
      return fail(XTBLOOM_STATUS_INVALID_ARGUMENT, "scc_mixer must be XTBLOOM_SCC_MIXER_MODIFIED_BROYDEN");
    if (options->scc_mixer_history < 1 || options->scc_mixer_history > 64)
      return fail(XTBLOOM_STATUS_INVALID_ARGUMENT, "scc_mixer_history must be in [1, 64]");
    if (!(options->scc_mixer_damping > 0.0) || !(options->scc_mixer_damping <= 1.0))
      return fail(XTBLOOM_STATUS_INVALID_ARGUMENT, "scc_mixer_damping must be in (0, 1]");
    if (options->determinism != XTBLOOM_DETERMINISM_DEFAULT &&
        options->determinism != XTBLOOM_DETERMINISM_REPRODUCIBLE)
      return fail(XTBLOOM_STATUS_INVALID_ARGUMENT,
                  "determinism must be XTBLOOM_DETERMINISM_DEFAULT or XTBLOOM_DETERMINISM_REPRODUCIBLE");
    if (options->reserved_v3 != 0)
      return fail(XTBLOOM_STATUS_INVALID_ARGUMENT, "compute options reserved_v3 field must be zero");
    ro.history = options->scc_mixer_history;
    ro.damping = options->scc_mixer_damping;
    ro.reproducible = options->determinism == XTBLOOM_DETERMINISM_REPRODUCIBLE;
  }
  if (result != nullptr && result->reserved != 0)
    return fail(XTBLOOM_STATUS_INVALID_ARGUMENT, "batch result reserved field must be zero");
  // result descriptor
  const bool want_e = flags & XTBLOOM_COMPUTE_ENERGY, want_f = flags & XTBLOOM_COMPUTE_FORCES;
  const bool want_q = flags & XTBLOOM_COMPUTE_ATOMIC_CHARGES;
  const bool want_pf = (flags & XTBLOOM_COMPUTE_POINT_CHARGE_FORCES) && P > 0;
  const bool want_d = flags & XTBLOOM_COMPUTE_DIPOLE_MOMENTS;
  xtbloom_status_t s;
  if (result != nullptr && (
      (s = check_output(result->energies, sizeof(double) * B, "energies", want_e)) ||
      (s = check_output(result->forces, sizeof(double) * 3 * N, "forces", want_f)) ||
      (s = check_output(result->atomic_charges, sizeof(double) * N, "atomic_charges", want_q)) ||
      (s = check_output(result->point_charge_forces, sizeof(double) * 3 * P, "point_charge_forces", want_pf)) ||
      (s = check_output(result->scc_iterations, sizeof(int32_t) * B, "scc_iterations", true)) ||
      (s = check_output(result->scc_converged, sizeof(uint8_t) * B, "scc_converged", true)) ||
      (s = check_output(result->per_system_status, sizeof(int32_t) * B, "per_system_status", true))))
    return s;
  if (want_d && result != nullptr) {
    if (result->struct_size < XTBLOOM_BATCH_RESULT_V2_SIZE)
      return fail(XTBLOOM_STATUS_INVALID_ARGUMENT, "dipole moments need an ABI-v2 batch result");
    if ((s = check_output(result->dipole_moments, sizeof(double) * 3 * B, "dipole_moments", true))) return s;
  }
  if (result != nullptr && result->struct_size >= XTBLOOM_BATCH_RESULT_V2_SIZE &&
      (result->quadrupole_moments.data || result->wiberg_orders.data || result->spin_populations.data))
    return fail(XTBLOOM_STATUS_INVALID_ARGUMENT, "reserved result outputs must be NULL");
  // unsupported optional inputs on this backend
  if (buffer_active(batch->atomic_potential_shifts) || buffer_active(batch->charge_response_offsets) ||
      buffer_active(batch->charge_response_matrix) || batch->total_charge_response_elements != 0)
    return fail(XTBLOOM_STATUS_NOT_SUPPORTED, "periodic b + A q embedding is not supported by the ROCm backend");
  if (batch->struct_size >= XTBLOOM_BATCH_V3_SIZE && batch->total_interactions != 0)
    return fail(XTBLOOM_STATUS_NOT_SUPPORTED, "interaction attachments are not supported by the ROCm backend");
  // stage inputs
  if ((s = stage_input(batch->atom_offsets, sizeof(int64_t) * (B + 1), "atom_offsets", st.atom_offsets, true)) ||
      (s = stage_input(batch->atomic_numbers, sizeof(int32_t) * N, "atomic_numbers", st.Z, true)) ||
      (s = stage_input(batch->positions, sizeof(double) * 3 * N, "positions", st.xyz, true)) ||
      (s = stage_input(batch->molecular_charges, sizeof(double) * B, "molecular_charges", st.charges, true)) ||
      (s = stage_input(batch->unpaired_electrons, sizeof(int32_t) * B, "unpaired_electrons", st.unpaired, true)))
    return s;
  if (batch->struct_size >= XTBLOOM_BATCH_V2_SIZE && buffer_active(batch->spin_channels)) {
    if ((s = stage_input(batch->spin_channels, sizeof(int32_t) * B, "spin_channels", st.spin, true))) return s;
    const int32_t* ch = reinterpret_cast<const int32_t*>(st.spin.data());
    for (int64_t i = 0; i < B; ++i) {
      if (ch[i] == 2) return fail(XTBLOOM_STATUS_NOT_SUPPORTED, "unrestricted spin channels are not supported by the ROCm backend");
      if (ch[i] != 1) return fail(XTBLOOM_STATUS_INVALID_ARGUMENT, "spin_channels must be 1 or 2");
    }
  }
  if (batch->struct_size >= XTBLOOM_BATCH_V4_SIZE &&
      (buffer_active(batch->cell_matrices) || buffer_active(batch->periodic_axes)))
    return fail(XTBLOOM_STATUS_NOT_IMPLEMENTED, "periodic batches are not implemented");
  if (P > 0) {
    if ((s = stage_input(batch->point_charge_offsets, sizeof(int64_t) * (B + 1), "point_charge_offsets", st.poff, true)) ||
        (s = stage_input(batch->point_charge_positions, sizeof(double) * 3 * P, "point_charge_positions", st.pxyz, true)) ||
        (s = stage_input(batch->point_charge_values, sizeof(double) * P, "point_charge_values", st.pq, true)) ||
        (s = stage_input(batch->point_charge_gammas, sizeof(double) * P, "point_charge_gammas", st.pg, true)))
      return s;
  }
  const int64_t* aoff = reinterpret_cast<const int64_t*>(st.atom_offsets.data());
  if (aoff[0] != 0 || aoff[B] != N) return fail(XTBLOOM_STATUS_INVALID_ARGUMENT, "atom_offsets must start at 0 and end at total_atoms");
  for (int64_t i = 0; i < B; ++i)
    if (aoff[i + 1] <= aoff[i]) return fail(XTBLOOM_STATUS_INVALID_ARGUMENT, "atom_offsets must be strictly increasing");
  const double* xyz = reinterpret_cast<const double*>(st.xyz.data());
  for (int64_t i = 0; i < 3 * N; ++i)
    if (!std::isfinite(xyz[i])) return fail(XTBLOOM_STATUS_INVALID_ARGUMENT, "positions contain NaN or infinity");
  const double* chg = reinterpret_cast<const double*>(st.charges.data());
  for (int64_t i = 0; i < B; ++i)
    if (!std::isfinite(chg[i])) return fail(XTBLOOM_STATUS_INVALID_ARGUMENT, "molecular charges must be finite");
  // coincident atoms make CN/ES/repulsion undefined (reference coordination.cpp:236-246).
  // r^2 < 1e-12 needs |dx| < 1e-6: sweep each system's atoms in x order and
  // compare only neighbours inside that window.
  {
    std::vector<int64_t> idx;
    for (int64_t sys = 0; sys < B; ++sys) {
      idx.resize(aoff[sys + 1] - aoff[sys]);
      std::iota(idx.begin(), idx.end(), aoff[sys]);
      std::sort(idx.begin(), idx.end(), [&](int64_t a, int64_t b) { return xyz[3 * a] < xyz[3 * b]; });
      for (size_t i = 0; i < idx.size(); ++i)
        for (size_t j = i + 1; j < idx.size() && xyz[3 * idx[j]] - xyz[3 * idx[i]] < 1e-6; ++j) {
          const int64_t a = idx[i], b = idx[j];
          const double dx = xyz[3 * a] - xyz[3 * b], dy = xyz[3 * a + 1] - xyz[3 * b + 1],
                       dz = xyz[3 * a + 2] - xyz[3 * b + 2];
          if (dx * dx + dy * dy + dz * dz < 1e-12)
            return fail(XTBLOOM_STATUS_INVALID_ARGUMENT, "coincident or near-coincident atoms");
        }
    }
  }
  if (P > 0) {
    const int64_t* po = reinterpret_cast<const int64_t*>(st.poff.data());
    if (po[0] != 0 || po[B] != P) return fail(XTBLOOM_STATUS_INVALID_ARGUMENT, "point_charge_offsets must start at 0 and end at total_point_charges");
    for (int64_t i = 0; i < B; ++i)
      if (po[i + 1] < po[i]) return fail(XTBLOOM_STATUS_INVALID_ARGUMENT, "point_charge_offsets must be monotone");
    const double* pxyz = reinterpret_cast<const double*>(st.pxyz.data());
    const double* pq = reinterpret_cast<const double*>(st.pq.data());
    const double* pg = reinterpret_cast<const double*>(st.pg.data());
    for (int64_t i = 0; i < P; ++i)
      if (!std::isfinite(pq[i]) || !(pg[i] > 0.0) || !std::isfinite(pg[i]) || !std::isfinite(pxyz[3 * i]) ||
          !std::isfinite(pxyz[3 * i + 1]) || !std::isfinite(pxyz[3 * i + 2]))
        return fail(XTBLOOM_STATUS_INVALID_ARGUMENT, "point charges must be finite with positive gammas");
  }
  hb.batch = B;
  hb.atoms = N;
  hb.points = P;
  hb.atom_offsets = aoff;
  hb.Z = reinterpret_cast<const int32_t*>(st.Z.data());
  hb.xyz = xyz;
  hb.charges = chg;
  hb.unpaired = reinterpret_cast<const int32_t*>(st.unpaired.data());
  if (P > 0) {
    hb.point_offsets = reinterpret_cast<const int64_t*>(st.poff.data());
    hb.pc_xyz = reinterpret_cast<const double*>(st.pxyz.data());
    hb.pc_q = reinterpret_cast<const double*>(st.pq.data());
    hb.pc_gamma = reinterpret_cast<const double*>(st.pg.data());
  }
  ro.forces = want_f || want_pf;
  return XTBLOOM_STATUS_SUCCESS;
}

CallIdentity make_identity(const Staged& st, const xtbloom_compute_options_t* options,
                           const xtb::RunOptions& ro, uint32_t flags) {
  CallIdentity id;
  id.atom_offsets = st.atom_offsets;
  id.Z = st.Z;
  id.charges = st.charges;
  id.unpaired = st.unpaired;
  id.point_offsets = st.poff;
  id.pc_gamma = st.pg;
  id.spin = st.spin;
  id.flags = flags;
  id.model = options->model;
  id.max_iterations = ro.max_iterations;
  id.history = ro.history;
  id.determinism = options->struct_size >= XTBLOOM_COMPUTE_OPTIONS_V3_SIZE ? options->determinism : 0;
  id.charge_tolerance = ro.charge_tolerance;
  id.energy_tolerance = ro.energy_tolerance;
  id.kt = ro.kt;
  id.damping = ro.damping;
  return id;
}

// Creates the context's extra lanes a batch of B systems runs on; returns
// the lanes to use (the context device first).
std::vector<xtb::Device*> lanes_for(xtbloom_context_t* context, int64_t B) {
  const int want = xtb::lanes_for_batch(B);
  while ((int)context->lanes.size() + 1 < want) {
    std::string lane_err;
    xtb::Device* d = xtb::device_create(xtb::device_id(context->device), nullptr, lane_err);
    if (d == nullptr) break;  // fewer lanes: the batch runs in larger slices
    context->lanes.push_back(d);
  }
  std::vector<xtb::Device*> lanes{context->device};
  for (int k = 0; k + 1 < want && k < (int)context->lanes.size(); ++k) lanes.push_back(context->lanes[k]);
  return lanes;
}

// Execution path of a call, from XTBLOOM_AMD_PATH: "plain" (the FP64 path
// with one host check per SCC round) or "batched" (graph-captured SCC rounds
// driven from the device and the mixed-precision products).
bool batched_path() {
  const char* env = getenv("XTBLOOM_AMD_PATH");
  return env != nullptr && std::strcmp(env, "batched") == 0;
}

// xtbloom_compute (plan == NULL) and xtbloom_plan_compute.  *admitted (if
// given) tells whether the call got past admission (validation, plan and WARM
// checks) before it returned.
xtbloom_status_t compute_impl(xtbloom_context_t* context, xtbloom_plan_t* plan, const xtbloom_batch_t* batch,
                              const xtbloom_compute_options_t* options, xtbloom_batch_result_t* result,
                              xtbloom_amd_extension_t* ext, bool* admitted = nullptr,
                              bool no_alloc = false) {
  if (admitted) *admitted = false;
  if (context == nullptr || context->device == nullptr)
    return fail(XTBLOOM_STATUS_INVALID_ARGUMENT, "context is NULL");
  std::lock_guard<std::mutex> guard(context->lock);
  Staged st;
  xtb::HostBatch hb;
  xtb::RunOptions ro;
  uint32_t flags = 0;
  bool warm = false;
  xtbloom_status_t s = validate_and_stage(batch, options, result, st, hb, ro, flags, warm);
  if (s != XTBLOOM_STATUS_SUCCESS) return s;
  if (ext != nullptr && !valid_header(ext, sizeof(xtbloom_amd_extension_t)))
    return fail(XTBLOOM_STATUS_INVALID_ARGUMENT, "extension descriptor is invalid");
  CallIdentity id = make_identity(st, options, ro, flags);
  if (plan != nullptr && !(id == plan->id))
    return fail(XTBLOOM_STATUS_INVALID_ARGUMENT,
                "the batch topology or compute policy does not match the fixed-topology plan");
  // strict WARM: the latest fully converged call with the same identity, never
  // a fallback to FRESH (reference xtbloom.h:502-523)
  WarmCheckpoint& cp = plan ? plan->warm : context->warm;
  std::vector<double> warm_state;
  if (warm) {
    if (!cp.valid || !(cp.id == id))
      return fail(XTBLOOM_STATUS_INVALID_ARGUMENT,
                  "WARM SCC start requires a previous fully converged call with identical topology and "
                  "compute policy on this context");
    warm_state.swap(cp.state);
    ro.warm_state = warm_state.data();
  }
  // an accepted call consumes the checkpoint whatever its start mode
  if (admitted) *admitted = true;
  cp.valid = false;
  cp.state.clear();
  ro.no_alloc = no_alloc;
  ro.batched = batched_path();
  {
    const char* env = getenv("XTBLOOM_AMD_SPLIT32");
    ro.split32 = env != nullptr && std::strcmp(env, "1") == 0;
  }
  const int64_t B = hb.batch, N = hb.atoms, P = hb.points;
  std::vector<int32_t> state_dim(B);
  {
    std::string dim_err;
    if (xtb::batch_matrix_elements(hb, nullptr, dim_err, nullptr, state_dim.data()) < 0)
      return fail(XTBLOOM_STATUS_INVALID_ARGUMENT, dim_err);
  }
  int64_t nstate = 0;
  for (int32_t d : state_dim) nstate += d;
  std::vector<double> scc_state((size_t)nstate);
  std::vector<double> energies(B), grad(3 * N), charges(N), pcgrad(3 * P), dipoles(3 * B);
  std::vector<int32_t> iters(B), status(B);
  std::vector<uint8_t> conv(B);
  xtb::HostOutputs out;
  out.energies = energies.data();
  out.gradient = grad.data();
  out.charges = charges.data();
  out.pc_gradient = P > 0 ? pcgrad.data() : nullptr;
  out.dipoles = dipoles.data();
  out.iterations = iters.data();
  out.converged = conv.data();
  out.status = status.data();
  if (ext) {
    out.pc_dedq = ext->point_charge_potential_gradient;
    out.dbg_overlap = ext->overlap;
    out.dbg_h0 = ext->core_hamiltonian;
    out.dbg_density = ext->density;
    out.dbg_cn = ext->coordination_numbers;
    out.dbg_levels = ext->orbital_energies;
    out.dbg_occ = ext->occupations;
    out.dbg_components = ext->energy_components;
    out.kernel_stats = reinterpret_cast<xtb::KernelStat*>(ext->kernel_stats);
    out.kernel_stats_mask = ext->kernel_stats_mask;
  }
  out.scc_state = scc_state.data();
  std::string err;
  xtb::RunTiming timing;
  int rc = 0;
  try {
    const std::vector<xtb::TopoCache*>* caches = nullptr;
    if (plan) {
      caches = &plan->caches;
    } else {
      // the context's own topology cache: emptied when the topology changes
      if (context->topo.empty())
        for (int k = 0; k < 4; ++k) context->topo.push_back(xtb::topo_cache_create());
      if (!same_topology(context->topo_id, id)) {
        for (xtb::TopoCache* c : context->topo) xtb::topo_cache_retag(c, -1);
        context->topo_id = id;
      }
      caches = &context->topo;
    }
    // a no_alloc call runs on the context's own stream only (more lanes would
    // create devices, which allocates)
    rc = xtb::run_batch_lanes(no_alloc ? std::vector<xtb::Device*>{context->device} : lanes_for(context, B),
                              hb, ro, out, &timing, err, caches);
  } catch (const std::bad_alloc&) {
    return fail(XTBLOOM_STATUS_ALLOCATION_FAILED, "host allocation failed");
  } catch (...) {
    return fail(XTBLOOM_STATUS_INTERNAL_ERROR, "unexpected exception in the ROCm executor");
  }
  if (rc != 0) return fail(rc, err);
  if (std::all_of(status.begin(), status.end(), [](int32_t v) { return v == 0; })) {
    cp.id = std::move(id);
    cp.state = std::move(scc_state);
    cp.valid = true;
  }
  if (ext) {
    ext->setup_ms = timing.setup_ms;
    ext->scc_ms = timing.scc_ms;
    ext->force_ms = timing.force_ms;
    ext->total_ms = timing.total_ms;
    ext->scc_rounds = timing.scc_rounds;
  }
  // forces are -dE/dR
  for (auto& v : grad) v = -v;
  for (auto& v : pcgrad) v = -v;
  if ((s = publish(result->energies, energies.data(), (flags & XTBLOOM_COMPUTE_ENERGY) ? sizeof(double) * B : 0)) ||
      (s = publish(result->forces, grad.data(), (flags & XTBLOOM_COMPUTE_FORCES) ? sizeof(double) * 3 * N : 0)) ||
      (s = publish(result->atomic_charges, charges.data(), (flags & XTBLOOM_COMPUTE_ATOMIC_CHARGES) ? sizeof(double) * N : 0)) ||
      (s = publish(result->point_charge_forces, pcgrad.data(),
                   (flags & XTBLOOM_COMPUTE_POINT_CHARGE_FORCES) ? sizeof(double) * 3 * P : 0)) ||
      (s = publish(result->scc_iterations, iters.data(), sizeof(int32_t) * B)) ||
      (s = publish(result->scc_converged, conv.data(), sizeof(uint8_t) * B)) ||
      (s = publish(result->per_system_status, status.data(), sizeof(int32_t) * B)))
    return s;
  uint32_t rflags = 0;
  if ((flags & XTBLOOM_COMPUTE_DIPOLE_MOMENTS) && result->struct_size >= XTBLOOM_BATCH_RESULT_V2_SIZE) {
    if ((s = publish(result->dipole_moments, dipoles.data(), sizeof(double) * 3 * B))) return s;
    rflags |= XTBLOOM_RESULT_DIPOLE_MOMENTS;
  }
  result->flags = rflags;
  g_last_error.clear();
  return XTBLOOM_STATUS_SUCCESS;
}

}  // namespace

extern "C" {

const char* xtbloom_version_string(void) { return XTBLOOM_VERSION_STRING; }

const char* xtbloom_status_string(xtbloom_status_t status) {
  switch (status) {
    case XTBLOOM_STATUS_SUCCESS: return "success";
    case XTBLOOM_STATUS_INVALID_ARGUMENT: return "invalid argument";
    case XTBLOOM_STATUS_BACKEND_UNAVAILABLE: return "backend unavailable";
    case XTBLOOM_STATUS_NOT_SUPPORTED: return "not supported";
    case XTBLOOM_STATUS_ALLOCATION_FAILED: return "allocation failed";
    case XTBLOOM_STATUS_NOT_IMPLEMENTED: return "not implemented";
    case XTBLOOM_STATUS_INTERNAL_ERROR: return "internal error";
    case XTBLOOM_STATUS_SCC_NOT_CONVERGED: return "SCC not converged";
    case XTBLOOM_STATUS_EIGENSOLVER_FAILED: return "eigensolver failed";
  }
  return "unknown status";
}

const char* xtbloom_get_last_error(void) { return g_last_error.c_str(); }

xtbloom_status_t xtbloom_context_options_init(xtbloom_context_options_t* o, size_t size) {
  xtbloom_status_t s = init_struct(o, size, XTBLOOM_CONTEXT_OPTIONS_V1_SIZE, "context options");
  if (s != XTBLOOM_STATUS_SUCCESS) return s;
  o->backend = XTBLOOM_BACKEND_AUTO;
  o->device_id = -1;
  return s;
}

xtbloom_status_t xtbloom_batch_init(xtbloom_batch_t* b, size_t size) {
  return init_struct(b, size, XTBLOOM_BATCH_V1_SIZE, "batch");
}

xtbloom_status_t xtbloom_compute_options_init(xtbloom_compute_options_t* o, size_t size) {
  xtbloom_status_t s = init_struct(o, size, XTBLOOM_COMPUTE_OPTIONS_V1_SIZE, "compute options");
  if (s != XTBLOOM_STATUS_SUCCESS) return s;
  o->model = XTBLOOM_MODEL_GFN2_XTB;
  o->flags = XTBLOOM_COMPUTE_ENERGY | XTBLOOM_COMPUTE_FORCES;
  o->max_scc_iterations = 250;
  o->charge_tolerance = 1.0e-6;
  o->energy_tolerance = 1.0e-8;
  o->electronic_temperature = XTBLOOM_DEFAULT_ELECTRONIC_TEMPERATURE;
  if (size >= XTBLOOM_COMPUTE_OPTIONS_V2_SIZE) o->scc_start_mode = XTBLOOM_SCC_START_FRESH;
  if (size >= XTBLOOM_COMPUTE_OPTIONS_V3_SIZE) {
    o->scc_mixer = XTBLOOM_SCC_MIXER_MODIFIED_BROYDEN;
    o->scc_mixer_history = 8;
    o->scc_mixer_damping = 0.4;
    o->determinism = XTBLOOM_DETERMINISM_DEFAULT;
  }
  return s;
}

xtbloom_status_t xtbloom_batch_result_init(xtbloom_batch_result_t* r, size_t size) {
  return init_struct(r, size, XTBLOOM_BATCH_RESULT_V1_SIZE, "batch result");
}

xtbloom_status_t xtbloom_workspace_query_init(xtbloom_workspace_query_t* q, size_t size) {
  return init_struct(q, size, XTBLOOM_WORKSPACE_QUERY_V1_SIZE, "workspace query");
}

xtbloom_status_t xtbloom_request_info_init(xtbloom_request_info_t* i, size_t size) {
  return init_struct(i, size, XTBLOOM_REQUEST_INFO_V1_SIZE, "request info");
}

xtbloom_status_t xtbloom_context_create(const xtbloom_context_options_t* options,
                                        xtbloom_context_t** context) {
  if (context == nullptr) return fail(XTBLOOM_STATUS_INVALID_ARGUMENT, "context output is NULL");
  *context = nullptr;
  if (!valid_header(options, XTBLOOM_CONTEXT_OPTIONS_V1_SIZE))
    return fail(XTBLOOM_STATUS_INVALID_ARGUMENT, "context options are NULL, too small, or have a different api_version");
  if (options->backend == XTBLOOM_BACKEND_CPU || options->backend == XTBLOOM_BACKEND_CUDA)
    return fail(XTBLOOM_STATUS_BACKEND_UNAVAILABLE, "this library provides the ROCm backend only");
  if (options->backend != XTBLOOM_BACKEND_AUTO && options->backend != XTBLOOM_BACKEND_ROCM)
    return fail(XTBLOOM_STATUS_INVALID_ARGUMENT, "unknown backend");
  if (options->reserved != 0) return fail(XTBLOOM_STATUS_INVALID_ARGUMENT, "context options reserved field must be zero");
  std::string err;
  try {
    xtbloom_context_t* ctx = new xtbloom_context_t();
    ctx->device = xtb::device_create(options->device_id, options->stream, err);
    if (ctx->device == nullptr) {
      delete ctx;
      return fail(XTBLOOM_STATUS_BACKEND_UNAVAILABLE, err);
    }
    *context = ctx;
  } catch (const std::bad_alloc&) {
    return fail(XTBLOOM_STATUS_ALLOCATION_FAILED, "context allocation failed");
  }
  g_last_error.clear();
  return XTBLOOM_STATUS_SUCCESS;
}

void xtbloom_context_destroy(xtbloom_context_t* context) {
  if (context == nullptr) return;
  for (xtb::Device* d : context->lanes) xtb::device_destroy(d);
  xtb::device_destroy(context->device);
  delete context;
}

xtbloom_backend_t xtbloom_context_get_backend(const xtbloom_context_t* context) {
  return context ? XTBLOOM_BACKEND_ROCM : XTBLOOM_BACKEND_AUTO;
}

int32_t xtbloom_context_get_device_id(const xtbloom_context_t* context) {
  return context ? xtb::device_id(context->device) : -1;
}

xtbloom_status_t xtbloom_compute(xtbloom_context_t* context, const xtbloom_batch_t* batch,
                                 const xtbloom_compute_options_t* options,
                                 xtbloom_batch_result_t* result) {
  return compute_impl(context, nullptr, batch, options, result, nullptr);
}

xtbloom_status_t xtbloom_amd_compute_extended(xtbloom_context_t* context, const xtbloom_batch_t* batch,
                                              const xtbloom_compute_options_t* options,
                                              xtbloom_batch_result_t* result,
                                              xtbloom_amd_extension_t* extension) {
  return compute_impl(context, nullptr, batch, options, result, extension);
}

xtbloom_status_t xtbloom_amd_compute_no_alloc(xtbloom_context_t* context, const xtbloom_batch_t* batch,
                                              const xtbloom_compute_options_t* options,
                                              xtbloom_batch_result_t* result) {
  return compute_impl(context, nullptr, batch, options, result, nullptr, nullptr, true);
}

int64_t xtbloom_amd_packed_matrix_elements(const xtbloom_batch_t* batch) {
  if (!valid_header(batch, XTBLOOM_BATCH_V1_SIZE) || batch->batch_size <= 0 ||
      batch->atom_offsets.memory_space != XTBLOOM_MEMORY_HOST ||
      batch->atomic_numbers.memory_space != XTBLOOM_MEMORY_HOST || batch->atom_offsets.data == nullptr ||
      batch->atomic_numbers.data == nullptr)
    return -1;
  xtb::HostBatch hb;
  hb.batch = batch->batch_size;
  hb.atom_offsets = static_cast<const int64_t*>(batch->atom_offsets.data);
  hb.Z = static_cast<const int32_t*>(batch->atomic_numbers.data);
  std::string err;
  return xtb::batch_matrix_elements(hb, nullptr, err);
}

// ---- asynchronous requests, plans and result owners: not provided by this backend ----
static xtbloom_status_t not_supported(const char* what) {
  return fail(XTBLOOM_STATUS_NOT_SUPPORTED, std::string(what) + " is not supported by the ROCm backend");
}

// ---- requests: submissions settled inline (reference xtbloom.h:836-937) ----
xtbloom_status_t xtbloom_request_create(xtbloom_context_t* context, xtbloom_request_t** request) {
  if (request == nullptr) return fail(XTBLOOM_STATUS_INVALID_ARGUMENT, "request output is NULL");
  *request = nullptr;
  if (context == nullptr || context->device == nullptr)
    return fail(XTBLOOM_STATUS_INVALID_ARGUMENT, "context is NULL");
  try {
    xtbloom_request_t* r = new xtbloom_request_t();
    r->context = context;
    *request = r;
  } catch (const std::bad_alloc&) {
    return fail(XTBLOOM_STATUS_ALLOCATION_FAILED, "request allocation failed");
  }
  g_last_error.clear();
  return XTBLOOM_STATUS_SUCCESS;
}

static xtbloom_status_t request_info(xtbloom_request_t* request, xtbloom_request_info_t* info) {
  if (request == nullptr) return fail(XTBLOOM_STATUS_INVALID_ARGUMENT, "request is NULL");
  if (!valid_header(info, XTBLOOM_REQUEST_INFO_V1_SIZE) || info->reserved != 0)
    return fail(XTBLOOM_STATUS_INVALID_ARGUMENT, "request info is NULL, too small, or has a different api_version");
  info->state = request->state;
  info->completion_status = request->state == XTBLOOM_REQUEST_COMPLETE ? request->completion_status
                                                                      : XTBLOOM_STATUS_SUCCESS;
  info->result_flags = request->state == XTBLOOM_REQUEST_COMPLETE ? request->result_flags : 0u;
  g_last_error.clear();
  return XTBLOOM_STATUS_SUCCESS;
}

// every accepted submission is already COMPLETE: query and wait coincide
xtbloom_status_t xtbloom_request_query(xtbloom_request_t* request, xtbloom_request_info_t* info) {
  return request_info(request, info);
}
xtbloom_status_t xtbloom_request_wait(xtbloom_request_t* request, xtbloom_request_info_t* info) {
  return request_info(request, info);
}

const char* xtbloom_request_get_error(const xtbloom_request_t* request) {
  if (request == nullptr) {
    fail(XTBLOOM_STATUS_INVALID_ARGUMENT, "request is NULL");
    return nullptr;
  }
  return request->state == XTBLOOM_REQUEST_COMPLETE ? request->error.c_str() : "";
}

void xtbloom_request_destroy(xtbloom_request_t* request) { delete request; }

// Runs one submission inline: admission errors return directly and leave the
// request and every output untouched; an admitted call completes the request
// with its status.  The caller's result descriptor is copied and never
// modified: completed result flags go to the request.
static xtbloom_status_t enqueue_impl(xtbloom_context_t* context, xtbloom_plan_t* plan,
                                     const xtbloom_batch_t* batch, const xtbloom_compute_options_t* options,
                                     const xtbloom_batch_result_t* result, xtbloom_request_t* request) {
  if (request == nullptr) return fail(XTBLOOM_STATUS_INVALID_ARGUMENT, "request is NULL");
  if (request->context != context)
    return fail(XTBLOOM_STATUS_INVALID_ARGUMENT, "request belongs to another context");
  if (!valid_header(result, XTBLOOM_BATCH_RESULT_V1_SIZE))
    return fail(XTBLOOM_STATUS_INVALID_ARGUMENT, "batch result is NULL, too small, or has a different api_version");
  xtbloom_batch_result_t copy;
  std::memset(&copy, 0, sizeof(copy));
  std::memcpy(&copy, result, std::min<size_t>(sizeof(copy), result->struct_size));
  bool admitted = false;
  const xtbloom_status_t s = compute_impl(context, plan, batch, options, &copy, nullptr, &admitted);
  if (!admitted) return s;
  request->state = XTBLOOM_REQUEST_COMPLETE;
  request->completion_status = s;
  request->result_flags = s == XTBLOOM_STATUS_SUCCESS ? copy.flags : 0u;
  request->error = s == XTBLOOM_STATUS_SUCCESS ? std::string() : g_last_error;
  g_last_error.clear();
  return XTBLOOM_STATUS_SUCCESS;
}

xtbloom_status_t xtbloom_compute_enqueue(xtbloom_context_t* context, const xtbloom_batch_t* batch,
                                         const xtbloom_compute_options_t* options,
                                         const xtbloom_batch_result_t* result, xtbloom_request_t* request) {
  return enqueue_impl(context, nullptr, batch, options, result, request);
}

// ---- fixed-topology plans (reference xtbloom.h:939-991) ----
xtbloom_status_t xtbloom_plan_create(xtbloom_context_t* context, const xtbloom_batch_t* batch,
                                     const xtbloom_compute_options_t* options, xtbloom_plan_t** plan) {
  if (plan == nullptr) return fail(XTBLOOM_STATUS_INVALID_ARGUMENT, "plan output is NULL");
  *plan = nullptr;
  if (context == nullptr || context->device == nullptr)
    return fail(XTBLOOM_STATUS_INVALID_ARGUMENT, "context is NULL");
  std::lock_guard<std::mutex> guard(context->lock);
  Staged st;
  xtb::HostBatch hb;
  xtb::RunOptions ro;
  uint32_t flags = 0;
  bool warm = false;
  xtbloom_status_t s = validate_and_stage(batch, options, nullptr, st, hb, ro, flags, warm);
  if (s != XTBLOOM_STATUS_SUCCESS) return s;
  try {
    std::unique_ptr<xtbloom_plan_t> p(new xtbloom_plan_t());
    p->context = context;
    p->id = make_identity(st, options, ro, flags);
    for (int k = 0; k < 4; ++k) p->caches.push_back(xtb::topo_cache_create());
    // reserve: host topology of every lane slice and the device workspace
    xtb::RunOptions reserve = ro;
    reserve.reserve_only = true;
    xtb::HostOutputs none;
    xtb::RunTiming timing;
    std::string err;
    const int rc = xtb::run_batch_lanes(lanes_for(context, hb.batch), hb, reserve, none, &timing, err,
                                        &p->caches);
    if (rc != 0) return fail(rc, err);
    std::vector<int32_t> dims(hb.batch);
    if (xtb::batch_matrix_elements(hb, nullptr, err, nullptr, dims.data()) < 0)
      return fail(XTBLOOM_STATUS_INVALID_ARGUMENT, err);
    int64_t nstate = 0;
    for (int32_t d : dims) nstate += d;
    p->device_bytes = timing.workspace_bytes;
    // per call: the staged inputs, the host outputs and the SCC state (twice:
    // the call's and the checkpoint)
    p->host_bytes = (uint64_t)(st.atom_offsets.size() + st.Z.size() + st.xyz.size() + st.charges.size() +
                               st.unpaired.size() + st.poff.size() + st.pxyz.size() + st.pq.size() +
                               st.pg.size() + st.spin.size()) +
                    sizeof(double) * (uint64_t)(5 * hb.batch + 4 * hb.atoms + 4 * hb.points + 2 * nstate) +
                    (sizeof(int32_t) * 2 + 1) * (uint64_t)hb.batch;
    *plan = p.release();
  } catch (const std::bad_alloc&) {
    return fail(XTBLOOM_STATUS_ALLOCATION_FAILED, "plan allocation failed");
  }
  g_last_error.clear();
  return XTBLOOM_STATUS_SUCCESS;
}

void xtbloom_plan_destroy(xtbloom_plan_t* plan) { delete plan; }

xtbloom_status_t xtbloom_plan_query_workspace(const xtbloom_plan_t* plan, xtbloom_workspace_query_t* query) {
  if (plan == nullptr) return fail(XTBLOOM_STATUS_INVALID_ARGUMENT, "plan is NULL");
  if (!valid_header(query, XTBLOOM_WORKSPACE_QUERY_V1_SIZE))
    return fail(XTBLOOM_STATUS_INVALID_ARGUMENT, "workspace query is NULL, too small, or has a different api_version");
  if (query->compute_flags != plan->id.flags)
    return fail(XTBLOOM_STATUS_INVALID_ARGUMENT, "compute_flags must equal the plan's compute flags");
  query->host_required_bytes = plan->host_bytes;
  query->host_required_alignment = alignof(std::max_align_t);
  query->device_required_bytes = plan->device_bytes;
  query->device_required_alignment = 256;
  g_last_error.clear();
  return XTBLOOM_STATUS_SUCCESS;
}

xtbloom_status_t xtbloom_plan_compute(xtbloom_plan_t* plan, const xtbloom_batch_t* batch,
                                      const xtbloom_compute_options_t* options, xtbloom_batch_result_t* result) {
  if (plan == nullptr) return fail(XTBLOOM_STATUS_INVALID_ARGUMENT, "plan is NULL");
  return compute_impl(plan->context, plan, batch, options, result, nullptr);
}
xtbloom_status_t xtbloom_plan_compute_enqueue(xtbloom_plan_t* plan, const xtbloom_batch_t* batch,
                                              const xtbloom_compute_options_t* options,
                                              const xtbloom_batch_result_t* result, xtbloom_request_t* request) {
  if (plan == nullptr) return fail(XTBLOOM_STATUS_INVALID_ARGUMENT, "plan is NULL");
  return enqueue_impl(plan->context, plan, batch, options, result, request);
}
xtbloom_status_t xtbloom_result_owner_options_init(xtbloom_result_owner_options_t* o, size_t size) {
  xtbloom_status_t s = init_struct(o, size, XTBLOOM_RESULT_OWNER_OPTIONS_V1_SIZE, "result owner options");
  if (s == XTBLOOM_STATUS_SUCCESS) o->device_id = -1;
  return s;
}
xtbloom_status_t xtbloom_result_owner_create(const xtbloom_result_owner_options_t*,
                                             xtbloom_result_owner_t** owner) {
  if (owner) *owner = nullptr;
  return not_supported("DLPack result owner");
}
xtbloom_status_t xtbloom_result_owner_buffer(const xtbloom_result_owner_t*, xtbloom_buffer_t*) {
  return not_supported("DLPack result owner");
}
void xtbloom_result_owner_retain(xtbloom_result_owner_t*) {}
void xtbloom_result_owner_release(xtbloom_result_owner_t*) {}
xtbloom_status_t xtbloom_result_owner_export_dltensor(const xtbloom_result_owner_t*,
                                                      const xtbloom_dlpack_view_t*, int, void**) {
  return not_supported("DLPack result owner");
}

}  // extern "C"
