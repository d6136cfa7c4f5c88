// xtbloom C ABI served by the MI355X GFN2 executor.
//
// Entry-point semantics follow the reference ABI (src/api.cpp, include/xtbloom/
// xtbloom.h): versioned struct prefixes, thread-local last error, validation of
// every descriptor before any output byte is written, per-system numerical
// failures reported as data (SUCCESS + NaN slices + per_system_status), and no
// exception crosses the boundary.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstdlib>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <limits>
#include <memory>
#include <mutex>
#include <new>
#include <numeric>
#include <string>
#include <thread>
#include <utility>
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
// positions, point-charge values) is not part of it.  The uniform-field
// attachments (presence per system and the field vectors, reference
// gfn2_cpu_execution.cpp:564-597) belong to the strict WARM identity only: a
// plan accepts another field, as it accepts another geometry
// (gfn2_cpu_execution.cpp:657-671).
struct CallIdentity {
  std::vector<char> atom_offsets, Z, charges, unpaired, point_offsets, pc_gamma, spin;
  uint32_t flags = 0;
  int32_t model = 0, max_iterations = 0, history = 0, determinism = 0;
  double charge_tolerance = 0.0, energy_tolerance = 0.0, kt = 0.0, damping = 0.0;
  // empty when the call attaches no field, else one entry (three) per system;
  // the vectors compare by value, as the reference's key does
  std::vector<uint8_t> field_on;
  std::vector<double> field;
  // the external charge operator's presence (its b and A are numerical
  // inputs, gfn2_cpu_execution.cpp:627-643)
  bool charge_operator = false;
  bool same_plan(const CallIdentity& o) const {
    // doubles compare by their bits: the identity is exact
    auto same = [](double a, double b) { return std::memcmp(&a, &b, sizeof(double)) == 0; };
    return atom_offsets == o.atom_offsets && Z == o.Z && charges == o.charges && unpaired == o.unpaired &&
           point_offsets == o.point_offsets && pc_gamma == o.pc_gamma && spin == o.spin && flags == o.flags &&
           model == o.model && max_iterations == o.max_iterations && history == o.history &&
           determinism == o.determinism && same(charge_tolerance, o.charge_tolerance) &&
           same(energy_tolerance, o.energy_tolerance) && same(kt, o.kt) && same(damping, o.damping) &&
           charge_operator == o.charge_operator;
  }
  bool operator==(const CallIdentity& o) const {
    return same_plan(o) && field_on == o.field_on && field == o.field;
  }
  // the first part of the identity that differs from o (for diagnostics)
  const char* difference(const CallIdentity& o) const {
    auto same = [](double a, double b) { return std::memcmp(&a, &b, sizeof(double)) == 0; };
    if (atom_offsets != o.atom_offsets || Z != o.Z) return "atoms";
    if (charges != o.charges || unpaired != o.unpaired) return "molecular charges or unpaired electrons";
    if (point_offsets != o.point_offsets || pc_gamma != o.pc_gamma) return "point-charge layout";
    if (spin != o.spin) return "spin channels";
    if (charge_operator != o.charge_operator) return "external charge operator";
    if (field_on != o.field_on || field != o.field) return "electric field";
    if (flags != o.flags || model != o.model || determinism != o.determinism) return "flags, model or determinism";
    if (max_iterations != o.max_iterations || history != o.history || !same(charge_tolerance, o.charge_tolerance) ||
        !same(energy_tolerance, o.energy_tolerance) || !same(kt, o.kt) || !same(damping, o.damping))
      return "SCC policy";
    return "nothing";
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

struct Job;

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
  // Asynchronous submissions (reference xtbloom.h:892-925): admitted jobs run
  // in submission order on one worker thread; synchronous calls wait until the
  // queue has drained, so every call on a context keeps submission order.
  // queue, busy and stop are guarded by `lock`.
  std::condition_variable cv;
  std::deque<Job*> queue;
  bool busy = false, stop = false;
  std::thread worker;
};

// Asynchronous submission owner (reference xtbloom.h:41-47, 836-878):
// IDLE after creation, PENDING from an accepted enqueue until the context's
// worker has run the call and published its outputs, then COMPLETE with the
// call's status, result flags and diagnostic.  Admission errors return from
// enqueue directly and leave the request unchanged.  The fields are guarded by
// `m`; the worker completes the request and notifies under it.
struct xtbloom_request {
  xtbloom_context_t* context = nullptr;
  int32_t state = XTBLOOM_REQUEST_IDLE;
  int32_t completion_status = XTBLOOM_STATUS_SUCCESS;
  uint32_t result_flags = 0;
  std::string error;
  std::mutex m;
  std::condition_variable cv;
};

// The inputs the packed host topology depends on: atom offsets, element
// numbers, molecular charges, unpaired electrons, point-charge offsets and
// gammas, spin channels.
bool same_topology(const CallIdentity& a, const CallIdentity& b) {
  return a.atom_offsets == b.atom_offsets && a.Z == b.Z && a.charges == b.charges &&
         a.unpaired == b.unpaired && a.point_offsets == b.point_offsets && a.pc_gamma == b.pc_gamma &&
         a.spin == b.spin;
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

// A result arena (reference xtbloom.h:1011-1145): host memory or ROCm device
// memory, freed when the last reference (the creator's, an explicit retain,
// or one per exported tensor) is released.
struct xtbloom_result_owner {
  std::atomic<int64_t> refs{1};
  xtbloom_memory_space_t memory_space = XTBLOOM_MEMORY_HOST;
  int32_t device_id = -1;
  uint64_t size = 0;
  void* data = nullptr;
  ~xtbloom_result_owner();
};

// DLPack 1.0 managed-tensor layouts, written from the published DLPack
// specification (dlpack.h v1.0: DLDevice, DLDataType, DLTensor,
// DLManagedTensor, DLManagedTensorVersioned); the byte layout is the
// interchange format, so it must be exactly this.
enum { kDLCPU = 1, kDLROCM = 10 };
enum { kDLInt = 0, kDLUInt = 1, kDLFloat = 2 };
struct DLTensorLayout {
  void* data;
  struct {
    int32_t device_type;
    int32_t device_id;
  } device;
  int32_t ndim;
  struct {
    uint8_t code;
    uint8_t bits;
    uint16_t lanes;
  } dtype;
  int64_t* shape;
  int64_t* strides;
  uint64_t byte_offset;
};
struct DLManagedLegacy {
  DLTensorLayout dl_tensor;
  void* manager_ctx;
  void (*deleter)(DLManagedLegacy*);
};
struct DLManagedVersioned {
  uint32_t version_major;
  uint32_t version_minor;
  void* manager_ctx;
  void (*deleter)(DLManagedVersioned*);
  uint64_t flags;
  DLTensorLayout dl_tensor;
};
static_assert(sizeof(DLTensorLayout) == 48 && sizeof(DLManagedLegacy) == 64 && sizeof(DLManagedVersioned) == 80,
              "DLPack 1.0 layouts");

// One exported tensor: its managed struct (one of the two forms), its shape
// storage and an arena reference, all released by the consumer's deleter.
struct ManagedExport {
  DLManagedLegacy legacy;
  DLManagedVersioned versioned;
  int64_t shape[XTBLOOM_DLPACK_MAX_NDIM];
  xtbloom_result_owner_t* owner = nullptr;
  static void delete_legacy(DLManagedLegacy* t);
  static void delete_versioned(DLManagedVersioned* t);
};

namespace {

thread_local std::string g_last_error;

xtbloom_status_t fail(xtbloom_status_t status, std::string message) {
  g_last_error = std::move(message);
  return status;
}

constexpr uint64_t kArenaAlign = 256;

// selects a device for the scope and restores the caller's on exit
struct DeviceScope {
  int previous = -1;
  bool ok = false;
  explicit DeviceScope(int device) {
    if (hipGetDevice(&previous) != hipSuccess) previous = -1;
    ok = hipSetDevice(device) == hipSuccess;
  }
  ~DeviceScope() {
    if (previous >= 0) (void)hipSetDevice(previous);
  }
};

}  // namespace

xtbloom_result_owner::~xtbloom_result_owner() {
  if (data == nullptr) return;
  if (memory_space == XTBLOOM_MEMORY_HOST) {
    std::free(data);
  } else {
    DeviceScope scope(device_id);
    (void)hipFree(data);
  }
}

void ManagedExport::delete_legacy(DLManagedLegacy* t) {
  ManagedExport* m = static_cast<ManagedExport*>(t->manager_ctx);
  xtbloom_result_owner_release(m->owner);
  delete m;
}

void ManagedExport::delete_versioned(DLManagedVersioned* t) {
  ManagedExport* m = static_cast<ManagedExport*>(t->manager_ctx);
  xtbloom_result_owner_release(m->owner);
  delete m;
}

namespace {

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
  // uniform-field attachments: empty without any, else [B] flags and [3B] vectors
  std::vector<uint8_t> field_on;
  std::vector<double> field;
  // external charge operator: b [N] and A (packed n_atoms^2 per system), each
  // empty when absent
  std::vector<double> pshift, presp;
};

// Native lattice descriptors (ABI-v4 suffix; reference xtbloom.h:186-200,
// validation.cpp:626-648, 1013-1089, lattice.hpp:85-123): cell_matrices and
// periodic_axes come together; a NONE item needs an all-zero cell, a mask
// with unknown bits is invalid, one- and two-dimensional masks are not
// supported, an XYZ cell must be finite, right-handed and nonsingular (the
// determinant of the rows, each scaled to unit length, above 2^-46).  Valid
// XYZ items are NOT_IMPLEMENTED; a batch whose items are all NONE runs as a
// molecular batch.
xtbloom_status_t check_lattice(const xtbloom_batch_t* batch, int64_t B) {
  if (batch->struct_size < XTBLOOM_BATCH_V4_SIZE) return XTBLOOM_STATUS_SUCCESS;
  const bool cells_on = buffer_active(batch->cell_matrices), axes_on = buffer_active(batch->periodic_axes);
  if (!cells_on && !axes_on) return XTBLOOM_STATUS_SUCCESS;
  if (!cells_on || !axes_on)
    return fail(XTBLOOM_STATUS_INVALID_ARGUMENT, "cell_matrices and periodic_axes must be supplied together");
  std::vector<char> cells, axes;
  xtbloom_status_t s;
  if ((s = stage_input(batch->cell_matrices, sizeof(double) * 9 * (size_t)B, "cell_matrices", cells, true)) ||
      (s = stage_input(batch->periodic_axes, sizeof(int32_t) * (size_t)B, "periodic_axes", axes, true)))
    return s;
  bool periodic = false;
  for (int64_t i = 0; i < B; ++i) {
    int32_t mask;
    double cell[9];
    std::memcpy(&mask, axes.data() + sizeof(int32_t) * i, sizeof(mask));
    std::memcpy(cell, cells.data() + sizeof(cell) * i, sizeof(cell));
    if (mask == XTBLOOM_PERIODIC_AXES_NONE) {
      for (double v : cell)
        if (v != 0.0) return fail(XTBLOOM_STATUS_INVALID_ARGUMENT, "a nonperiodic batch item must use an all-zero cell matrix");
      continue;
    }
    if ((mask & ~XTBLOOM_PERIODIC_AXES_XYZ) != 0)
      return fail(XTBLOOM_STATUS_INVALID_ARGUMENT, "periodic_axes contains unknown mask bits");
    if (mask != XTBLOOM_PERIODIC_AXES_XYZ)
      return fail(XTBLOOM_STATUS_NOT_SUPPORTED, "one- and two-dimensional periodic axes are reserved but not supported");
    for (double v : cell)
      if (!std::isfinite(v)) return fail(XTBLOOM_STATUS_INVALID_ARGUMENT, "a periodic cell matrix contains NaN or infinity");
    double u[9];
    bool ok = true;
    for (int r = 0; r < 3 && ok; ++r) {
      double m = 0.0;
      for (int c = 0; c < 3; ++c) m = std::max(m, std::fabs(cell[3 * r + c]));
      if (!(m > 0.0)) { ok = false; break; }
      double nn = 0.0;
      for (int c = 0; c < 3; ++c) {
        u[3 * r + c] = cell[3 * r + c] / m;
        nn += u[3 * r + c] * u[3 * r + c];
      }
      const double norm = std::sqrt(nn);
      for (int c = 0; c < 3; ++c) u[3 * r + c] /= norm;
    }
    const double det = ok ? u[0] * (u[4] * u[8] - u[5] * u[7]) + u[1] * (u[5] * u[6] - u[3] * u[8]) +
                                u[2] * (u[3] * u[7] - u[4] * u[6])
                          : 0.0;
    if (!(det > 0x1p-46) || !std::isfinite(det))
      return fail(XTBLOOM_STATUS_INVALID_ARGUMENT, "a periodic cell must be finite, right-handed, and nonsingular");
    periodic = true;
  }
  if (periodic)
    return fail(XTBLOOM_STATUS_NOT_IMPLEMENTED,
                "native lattice/PBC descriptors are valid but periodic GFN2 execution is not implemented");
  return XTBLOOM_STATUS_SUCCESS;
}

// External charge operator v = b + A q, E = q.b + q.A.q/2 (periodic QM/MM
// coupling; reference xtbloom.h:424-431, validation.cpp:401-403, 676-700,
// 1308-1360, gfn2_cpu_execution.cpp:448-465, 538-560): b is active when its
// buffer has data and bytes; A needs its count, offsets and matrix together,
// the offsets pack one square n_atoms^2 block per system in order, and every
// block is exactly symmetric.  All values finite.
xtbloom_status_t stage_charge_operator(const xtbloom_batch_t* batch, int64_t B, int64_t N,
                                       const int64_t* aoff, Staged& st) {
  xtbloom_status_t s;
  st.pshift.clear();
  st.presp.clear();
  std::vector<char> raw;
  if (batch->atomic_potential_shifts.data != nullptr && batch->atomic_potential_shifts.size_bytes != 0) {
    if ((s = stage_input(batch->atomic_potential_shifts, sizeof(double) * (size_t)N, "atomic_potential_shifts",
                         raw, true)))
      return s;
    st.pshift.resize((size_t)N);
    std::memcpy(st.pshift.data(), raw.data(), raw.size());
  }
  const int64_t total = batch->total_charge_response_elements;
  if (total < 0) return fail(XTBLOOM_STATUS_INVALID_ARGUMENT, "total_charge_response_elements must be nonnegative");
  const bool enabled = total != 0 || buffer_active(batch->charge_response_offsets) ||
                       buffer_active(batch->charge_response_matrix);
  if (enabled) {
    if (total == 0 || !buffer_active(batch->charge_response_offsets) || !buffer_active(batch->charge_response_matrix))
      return fail(XTBLOOM_STATUS_INVALID_ARGUMENT, "charge response count, offsets, and matrix must be supplied together");
    if (total > ((int64_t)1 << 40))
      return fail(XTBLOOM_STATUS_INVALID_ARGUMENT, "total_charge_response_elements overflows the addressable byte size");
    std::vector<char> offs;
    if ((s = stage_input(batch->charge_response_offsets, sizeof(int64_t) * (size_t)(B + 1), "charge_response_offsets",
                         offs, true)))
      return s;
    const int64_t* ro = reinterpret_cast<const int64_t*>(offs.data());
    if (ro[0] != 0) return fail(XTBLOOM_STATUS_INVALID_ARGUMENT, "charge_response_offsets must begin with zero");
    int64_t packed = 0;
    for (int64_t i = 0; i < B; ++i) {
      if (ro[i + 1] < ro[i])
        return fail(XTBLOOM_STATUS_INVALID_ARGUMENT, "charge_response_offsets must be monotonically nondecreasing");
      const int64_t n = aoff[i + 1] - aoff[i];
      if (ro[i + 1] - ro[i] != n * n)
        return fail(XTBLOOM_STATUS_INVALID_ARGUMENT, "charge_response_offsets do not pack one square matrix per system");
      packed += n * n;
    }
    if (ro[B] != total || packed != total)
      return fail(XTBLOOM_STATUS_INVALID_ARGUMENT, "total_charge_response_elements does not match the packed square matrices");
    if ((s = stage_input(batch->charge_response_matrix, sizeof(double) * (size_t)total, "charge_response_matrix", raw,
                         true)))
      return s;
    st.presp.resize((size_t)total);
    std::memcpy(st.presp.data(), raw.data(), raw.size());
  }
  for (double v : st.pshift)
    if (!std::isfinite(v)) return fail(XTBLOOM_STATUS_INVALID_ARGUMENT, "periodic b/A inputs contain NaN or infinity");
  for (double v : st.presp)
    if (!std::isfinite(v)) return fail(XTBLOOM_STATUS_INVALID_ARGUMENT, "periodic b/A inputs contain NaN or infinity");
  if (!st.presp.empty()) {
    int64_t base = 0;
    for (int64_t i = 0; i < B; ++i) {
      const int64_t n = aoff[i + 1] - aoff[i];
      const double* A = st.presp.data() + base;
      for (int64_t r = 0; r < n; ++r)
        for (int64_t c = r + 1; c < n; ++c)
          if (A[r * n + c] != A[c * n + r])
            return fail(XTBLOOM_STATUS_INVALID_ARGUMENT, "charge_response_matrix must be exactly symmetric per system");
      base += n * n;
    }
  }
  return XTBLOOM_STATUS_SUCCESS;
}

bool known_interaction(int32_t type) {
  switch (type) {
    case XTBLOOM_INTERACTION_ELECTRIC_FIELD:
    case XTBLOOM_INTERACTION_ELECTRIC_FIELD_GRADIENT:
    case XTBLOOM_INTERACTION_POINT_CHARGES_MULTIPOLE:
    case XTBLOOM_INTERACTION_ATOMIC_POTENTIAL_GRID:
    case XTBLOOM_INTERACTION_ALPB_SOLVATION:
    case XTBLOOM_INTERACTION_GBSA_SOLVATION:
    case XTBLOOM_INTERACTION_GB_SOLVATION:
    case XTBLOOM_INTERACTION_GBE_SOLVATION:
    case XTBLOOM_INTERACTION_DDX_SOLVATION:
    case XTBLOOM_INTERACTION_D3_DISPERSION:
    case XTBLOOM_INTERACTION_D4_VARIANT_DISPERSION:
    case XTBLOOM_INTERACTION_HALOGEN_BOND:
      return true;
    default:
      return false;
  }
}

// ABI-v3 interaction attachments (reference xtbloom.h:160-222, validation.cpp:
// 552-625, 920-1010, 1115-1140): every descriptor is checked structurally
// (flags zero, a known tag other than NONE, system index in the batch, the
// block inside the payload, an aligned block_version header), the uniform
// electric field block exactly (32 bytes, 8-byte aligned, block_version 1,
// reserved zero, finite components), at most one attachment per (system,
// tag); only then do the reserved tags return NOT_IMPLEMENTED.  Descriptors
// and payload may be host or ROCm device memory; they are staged once.
xtbloom_status_t stage_interactions(const xtbloom_batch_t* batch, int64_t B, Staged& st) {
  st.field_on.clear();
  st.field.clear();
  if (batch->struct_size < XTBLOOM_BATCH_V3_SIZE) return XTBLOOM_STATUS_SUCCESS;
  const int64_t count = batch->total_interactions;
  if (count < 0) return fail(XTBLOOM_STATUS_INVALID_ARGUMENT, "total_interactions must be nonnegative");
  if (count == 0) return XTBLOOM_STATUS_SUCCESS;
  if (count > ((int64_t)1 << 40) / (int64_t)sizeof(xtbloom_interaction_t))
    return fail(XTBLOOM_STATUS_INVALID_ARGUMENT, "total_interactions overflows the addressable byte size");
  std::vector<char> desc, payload;
  xtbloom_status_t s;
  if ((s = stage_input(batch->interaction_descriptors, sizeof(xtbloom_interaction_t) * (size_t)count,
                       "interaction_descriptors", desc, true)))
    return s;
  if (!buffer_active(batch->interaction_payload))
    return fail(XTBLOOM_STATUS_INVALID_ARGUMENT, "interaction_payload is required when interactions are present");
  const uint64_t pbytes = batch->interaction_payload.size_bytes;
  const uintptr_t pbase = reinterpret_cast<uintptr_t>(batch->interaction_payload.data);
  std::vector<xtbloom_interaction_t> items((size_t)count);
  std::memcpy(items.data(), desc.data(), desc.size());
  bool any_field = false, reserved_tag = false;
  for (const xtbloom_interaction_t& it : items) {
    if (it.flags != 0) return fail(XTBLOOM_STATUS_INVALID_ARGUMENT, "interaction descriptor flags must be zero");
    if (it.type == XTBLOOM_INTERACTION_NONE || !known_interaction(it.type))
      return fail(XTBLOOM_STATUS_INVALID_ARGUMENT, "an interaction descriptor uses an unknown or NONE type tag");
    if (it.system_index < 0 || it.system_index >= B)
      return fail(XTBLOOM_STATUS_INVALID_ARGUMENT, "an interaction system_index lies outside the batch");
    if (it.payload_offset > pbytes || it.payload_size > pbytes - it.payload_offset)
      return fail(XTBLOOM_STATUS_INVALID_ARGUMENT, "an interaction payload block extends past interaction_payload");
    if (it.payload_offset > (uint64_t)(UINTPTR_MAX - pbase))
      return fail(XTBLOOM_STATUS_INVALID_ARGUMENT, "an interaction payload block address overflows uintptr_t");
    const uintptr_t addr = pbase + (uintptr_t)it.payload_offset;
    if (it.type != XTBLOOM_INTERACTION_ELECTRIC_FIELD) {
      if (it.payload_size < sizeof(int32_t) || addr % alignof(int32_t) != 0)
        return fail(XTBLOOM_STATUS_INVALID_ARGUMENT, "an interaction payload block must contain an aligned block_version");
      reserved_tag = true;
      continue;
    }
    if (it.payload_size < 32 || addr % alignof(double) != 0)
      return fail(XTBLOOM_STATUS_INVALID_ARGUMENT, "an electric-field interaction payload block is undersized or misaligned");
    if (it.payload_size > 32)
      return fail(XTBLOOM_STATUS_INVALID_ARGUMENT, "an electric-field interaction payload block exceeds the released contract");
    any_field = true;
  }
  // the field blocks' contents
  if (any_field) {
    if ((s = stage_input(batch->interaction_payload, (size_t)pbytes, "interaction_payload", payload, true))) return s;
    for (const xtbloom_interaction_t& it : items) {
      if (it.type != XTBLOOM_INTERACTION_ELECTRIC_FIELD) continue;
      const char* block = payload.data() + it.payload_offset;
      int32_t version = 0, reserved = 0;
      double E[3];
      std::memcpy(&version, block, sizeof(version));
      std::memcpy(&reserved, block + 4, sizeof(reserved));
      std::memcpy(E, block + 8, sizeof(E));
      if (version != 1)
        return fail(XTBLOOM_STATUS_INVALID_ARGUMENT, "an electric-field interaction uses an unsupported block_version");
      if (reserved != 0)
        return fail(XTBLOOM_STATUS_INVALID_ARGUMENT, "an electric-field interaction reserved payload field must be zero");
      if (!std::isfinite(E[0]) || !std::isfinite(E[1]) || !std::isfinite(E[2]))
        return fail(XTBLOOM_STATUS_INVALID_ARGUMENT, "an electric-field interaction contains NaN or infinity");
    }
  }
  // at most one attachment per (system, tag)
  {
    std::vector<std::pair<int64_t, int32_t>> keys;
    keys.reserve(items.size());
    for (const xtbloom_interaction_t& it : items) keys.emplace_back(it.system_index, it.type);
    std::sort(keys.begin(), keys.end());
    if (std::adjacent_find(keys.begin(), keys.end()) != keys.end())
      return fail(XTBLOOM_STATUS_INVALID_ARGUMENT, "the same system has two attachments of the same interaction type");
  }
  if (reserved_tag)
    return fail(XTBLOOM_STATUS_NOT_IMPLEMENTED,
                "an interaction attachment uses a tag whose execution is reserved but not implemented");
  st.field_on.assign((size_t)B, 0);
  st.field.assign(3 * (size_t)B, 0.0);
  for (const xtbloom_interaction_t& it : items) {
    std::memcpy(&st.field[3 * (size_t)it.system_index], payload.data() + it.payload_offset + 8, 3 * sizeof(double));
    st.field_on[(size_t)it.system_index] = 1;
  }
  return XTBLOOM_STATUS_SUCCESS;
}

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
    for (int64_t i = 0; i < B; ++i)
      if (ch[i] != 1 && ch[i] != 2) return fail(XTBLOOM_STATUS_INVALID_ARGUMENT, "spin_channels must be 1 or 2");
  } else {
    // an absent buffer means one restricted channel per system; the identity
    // of a call therefore does not depend on whether the ones were given
    st.spin.resize(sizeof(int32_t) * (size_t)B);
    const int32_t one = 1;
    for (int64_t i = 0; i < B; ++i) std::memcpy(st.spin.data() + sizeof(int32_t) * i, &one, sizeof(one));
  }
  if ((s = stage_interactions(batch, B, st))) return s;
  if ((s = check_lattice(batch, B))) return s;
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
  if ((s = stage_charge_operator(batch, B, N, aoff, st))) return s;
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
  hb.field = st.field.empty() ? nullptr : st.field.data();
  hb.spin = st.spin.empty() ? nullptr : reinterpret_cast<const int32_t*>(st.spin.data());
  hb.pshift = st.pshift.empty() ? nullptr : st.pshift.data();
  hb.presp = st.presp.empty() ? nullptr : st.presp.data();
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
  id.field_on = st.field_on;
  id.field = st.field;
  id.charge_operator = !st.pshift.empty() || !st.presp.empty();
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
}  // namespace

// One admitted call: its staged inputs, options and identity, the consumed
// WARM state, and a copy of the caller's result descriptor (the output
// buffers it names stay borrowed until the call completes).
struct Job {
  xtbloom_plan_t* plan = nullptr;
  Staged st;
  xtb::HostBatch hb;
  xtb::RunOptions ro;
  uint32_t flags = 0;
  bool warm = false;
  CallIdentity id;
  std::vector<double> warm_state;
  xtbloom_batch_result_t result;
  xtbloom_amd_extension_t* ext = nullptr;
  xtbloom_request_t* request = nullptr;
  bool all_converged = false;
  std::vector<double> scc_state;
};

namespace {

// Admission (context lock held): validation and staging of every input,
// the plan identity, and strict WARM (reference xtbloom.h:502-523: the latest
// fully converged call with the same identity, never a fallback to FRESH).
// An admitted call consumes the checkpoint whatever its start mode.
xtbloom_status_t admit(xtbloom_context_t* context, xtbloom_plan_t* plan, const xtbloom_batch_t* batch,
                       const xtbloom_compute_options_t* options, const xtbloom_batch_result_t* result,
                       xtbloom_amd_extension_t* ext, bool no_alloc, Job& job) {
  xtbloom_status_t s = validate_and_stage(batch, options, result, job.st, job.hb, job.ro, job.flags, job.warm);
  if (s != XTBLOOM_STATUS_SUCCESS) return s;
  if (ext != nullptr && !valid_header(ext, sizeof(xtbloom_amd_extension_t)))
    return fail(XTBLOOM_STATUS_INVALID_ARGUMENT, "extension descriptor is invalid");
  job.id = make_identity(job.st, options, job.ro, job.flags);
  if (plan != nullptr && !job.id.same_plan(plan->id))
    return fail(XTBLOOM_STATUS_INVALID_ARGUMENT,
                "the batch topology or compute policy does not match the fixed-topology plan");
  WarmCheckpoint& cp = plan ? plan->warm : context->warm;
  if (job.warm) {
    if (!cp.valid)
      return fail(XTBLOOM_STATUS_INVALID_ARGUMENT,
                  "WARM SCC start requires a previous fully converged call with identical topology and "
                  "compute policy on this context (there is none)");
    if (!(cp.id == job.id))
      return fail(XTBLOOM_STATUS_INVALID_ARGUMENT,
                  std::string("WARM SCC start requires a previous fully converged call with identical topology "
                              "and compute policy on this context (the ") +
                      job.id.difference(cp.id) + " differ)");
    job.warm_state.swap(cp.state);
  }
  cp.valid = false;
  cp.state.clear();
  job.plan = plan;
  job.ext = ext;
  job.ro.no_alloc = no_alloc;
  job.ro.batched = batched_path();
  {
    const char* env = getenv("XTBLOOM_AMD_SPLIT32");
    job.ro.split32 = env != nullptr && std::strcmp(env, "1") == 0;
  }
  std::memset(&job.result, 0, sizeof(job.result));
  std::memcpy(&job.result, result, std::min<size_t>(sizeof(job.result), result->struct_size));
  return XTBLOOM_STATUS_SUCCESS;
}

// Execution of an admitted call: the device pipeline, then publication into
// the caller's output buffers.  Runs with no other execution on the context
// (the worker, or a synchronous caller after the queue has drained).
xtbloom_status_t execute(xtbloom_context_t* context, Job& job) {
  job.ro.warm_state = job.warm ? job.warm_state.data() : nullptr;
  xtbloom_status_t s = XTBLOOM_STATUS_SUCCESS;
  const int64_t B = job.hb.batch, N = job.hb.atoms, P = job.hb.points;
  std::vector<int32_t> state_dim(B);
  {
    std::string dim_err;
    if (xtb::batch_matrix_elements(job.hb, nullptr, dim_err, nullptr, state_dim.data()) < 0)
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
  xtbloom_amd_extension_t* ext = job.ext;
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
    if (job.plan) {
      caches = &job.plan->caches;
    } else {
      // the context's own topology cache: emptied when the topology changes
      if (context->topo.empty())
        for (int k = 0; k < 4; ++k) context->topo.push_back(xtb::topo_cache_create());
      if (!same_topology(context->topo_id, job.id)) {
        for (xtb::TopoCache* c : context->topo) xtb::topo_cache_retag(c, -1);
        context->topo_id = job.id;
      }
      caches = &context->topo;
    }
    // a no_alloc call runs on the context's own stream only (more lanes would
    // create devices, which allocates)
    rc = xtb::run_batch_lanes(job.ro.no_alloc ? std::vector<xtb::Device*>{context->device} : lanes_for(context, B),
                              job.hb, job.ro, out, &timing, err, caches);
  } catch (const std::bad_alloc&) {
    return fail(XTBLOOM_STATUS_ALLOCATION_FAILED, "host allocation failed");
  } catch (...) {
    return fail(XTBLOOM_STATUS_INTERNAL_ERROR, "unexpected exception in the ROCm executor");
  }
  if (rc != 0) return fail(rc, err);
  // the SCC state becomes the WARM checkpoint when every system converged
  // (committed by the caller of execute, under the context lock)
  job.all_converged = std::all_of(status.begin(), status.end(), [](int32_t v) { return v == 0; });
  if (job.all_converged) job.scc_state = std::move(scc_state);
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
  if ((s = publish(job.result.energies, energies.data(), (job.flags & XTBLOOM_COMPUTE_ENERGY) ? sizeof(double) * B : 0)) ||
      (s = publish(job.result.forces, grad.data(), (job.flags & XTBLOOM_COMPUTE_FORCES) ? sizeof(double) * 3 * N : 0)) ||
      (s = publish(job.result.atomic_charges, charges.data(), (job.flags & XTBLOOM_COMPUTE_ATOMIC_CHARGES) ? sizeof(double) * N : 0)) ||
      (s = publish(job.result.point_charge_forces, pcgrad.data(),
                   (job.flags & XTBLOOM_COMPUTE_POINT_CHARGE_FORCES) ? sizeof(double) * 3 * P : 0)) ||
      (s = publish(job.result.scc_iterations, iters.data(), sizeof(int32_t) * B)) ||
      (s = publish(job.result.scc_converged, conv.data(), sizeof(uint8_t) * B)) ||
      (s = publish(job.result.per_system_status, status.data(), sizeof(int32_t) * B)))
    return s;
  uint32_t rflags = 0;
  if ((job.flags & XTBLOOM_COMPUTE_DIPOLE_MOMENTS) && job.result.struct_size >= XTBLOOM_BATCH_RESULT_V2_SIZE) {
    if ((s = publish(job.result.dipole_moments, dipoles.data(), sizeof(double) * 3 * B))) return s;
    rflags |= XTBLOOM_RESULT_DIPOLE_MOMENTS;
  }
  // forces leave out the coordinate derivatives of a caller's b and A
  // (reference gfn2_cpu_execution.cpp:1983-1989)
  if (job.id.charge_operator) rflags |= XTBLOOM_RESULT_FORCES_EXCLUDE_EXTERNAL_OPERATOR_DERIVATIVES;
  job.result.flags = rflags;
  g_last_error.clear();
  return XTBLOOM_STATUS_SUCCESS;
}


void checkpoint_commit(xtbloom_context_t* context, Job& job) {
  if (!job.all_converged) return;
  WarmCheckpoint& cp = job.plan ? job.plan->warm : context->warm;
  cp.id = job.id;
  cp.state = std::move(job.scc_state);
  cp.valid = true;
}

// waits (context lock held through `guard`) until no submission is queued or running
void drain(xtbloom_context_t* context, std::unique_lock<std::mutex>& guard) {
  context->cv.wait(guard, [context] { return context->queue.empty() && !context->busy; });
}

// A synchronous call: after the submissions queued before it, admission and
// execution under the context lock.
xtbloom_status_t compute_impl(xtbloom_context_t* context, xtbloom_plan_t* plan, const xtbloom_batch_t* batch,
                              const xtbloom_compute_options_t* options, xtbloom_batch_result_t* result,
                              xtbloom_amd_extension_t* ext, bool* admitted = nullptr,
                              bool no_alloc = false) {
  if (admitted) *admitted = false;
  if (context == nullptr || context->device == nullptr)
    return fail(XTBLOOM_STATUS_INVALID_ARGUMENT, "context is NULL");
  std::unique_lock<std::mutex> guard(context->lock);
  drain(context, guard);
  std::unique_ptr<Job> job;
  try {
    job.reset(new Job());
  } catch (const std::bad_alloc&) {
    return fail(XTBLOOM_STATUS_ALLOCATION_FAILED, "host allocation failed");
  }
  xtbloom_status_t s = admit(context, plan, batch, options, result, ext, no_alloc, *job);
  if (s != XTBLOOM_STATUS_SUCCESS) return s;
  if (admitted) *admitted = true;
  try {
    s = execute(context, *job);
  } catch (const std::bad_alloc&) {
    s = fail(XTBLOOM_STATUS_ALLOCATION_FAILED, "host allocation failed");
  }
  checkpoint_commit(context, *job);
  if (s == XTBLOOM_STATUS_SUCCESS) result->flags = job->result.flags;
  return s;
}

// The context's worker: runs queued submissions in order and completes their
// requests.
void worker_loop(xtbloom_context_t* context) {
  std::unique_lock<std::mutex> guard(context->lock);
  for (;;) {
    context->cv.wait(guard, [context] { return context->stop || !context->queue.empty(); });
    if (context->queue.empty()) return;  // stopped and drained
    std::unique_ptr<Job> job(context->queue.front());
    context->queue.pop_front();
    context->busy = true;
    guard.unlock();
    xtbloom_status_t s;
    try {
      s = execute(context, *job);
    } catch (const std::bad_alloc&) {
      s = fail(XTBLOOM_STATUS_ALLOCATION_FAILED, "host allocation failed");
    } catch (...) {
      s = fail(XTBLOOM_STATUS_INTERNAL_ERROR, "unexpected exception in the ROCm executor");
    }
    std::string err = s == XTBLOOM_STATUS_SUCCESS ? std::string() : g_last_error;
    g_last_error.clear();
    guard.lock();
    checkpoint_commit(context, *job);
    context->busy = false;
    xtbloom_request_t* r = job->request;
    {
      // completed and notified under the request's mutex: a destroy waiting
      // for completion cannot free the request before the notification
      std::lock_guard<std::mutex> rg(r->m);
      r->completion_status = s;
      r->result_flags = s == XTBLOOM_STATUS_SUCCESS ? job->result.flags : 0u;
      r->error.swap(err);
      r->state = XTBLOOM_REQUEST_COMPLETE;
      r->cv.notify_all();
    }
    context->cv.notify_all();
  }
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
  {
    // queued submissions complete, then the worker stops
    std::lock_guard<std::mutex> guard(context->lock);
    context->stop = true;
  }
  context->cv.notify_all();
  if (context->worker.joinable()) context->worker.join();
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

static xtbloom_status_t request_info(xtbloom_request_t* request, xtbloom_request_info_t* info, bool wait) {
  if (request == nullptr) return fail(XTBLOOM_STATUS_INVALID_ARGUMENT, "request is NULL");
  if (!valid_header(info, XTBLOOM_REQUEST_INFO_V1_SIZE) || info->reserved != 0)
    return fail(XTBLOOM_STATUS_INVALID_ARGUMENT, "request info is NULL, too small, or has a different api_version");
  std::unique_lock<std::mutex> rg(request->m);
  if (wait) request->cv.wait(rg, [request] { return request->state != XTBLOOM_REQUEST_PENDING; });
  info->state = request->state;
  info->completion_status = request->state == XTBLOOM_REQUEST_COMPLETE ? request->completion_status
                                                                         : XTBLOOM_STATUS_SUCCESS;
  info->result_flags = request->state == XTBLOOM_REQUEST_COMPLETE ? request->result_flags : 0u;
  g_last_error.clear();
  return XTBLOOM_STATUS_SUCCESS;
}

xtbloom_status_t xtbloom_request_query(xtbloom_request_t* request, xtbloom_request_info_t* info) {
  return request_info(request, info, false);
}
xtbloom_status_t xtbloom_request_wait(xtbloom_request_t* request, xtbloom_request_info_t* info) {
  return request_info(request, info, true);
}

const char* xtbloom_request_get_error(const xtbloom_request_t* request) {
  if (request == nullptr) {
    fail(XTBLOOM_STATUS_INVALID_ARGUMENT, "request is NULL");
    return nullptr;
  }
  xtbloom_request_t* r = const_cast<xtbloom_request_t*>(request);
  std::lock_guard<std::mutex> rg(r->m);
  return r->state == XTBLOOM_REQUEST_COMPLETE ? r->error.c_str() : "";
}

// a PENDING request is destroyed after its submission completes
void xtbloom_request_destroy(xtbloom_request_t* request) {
  if (request == nullptr) return;
  {
    std::unique_lock<std::mutex> rg(request->m);
    request->cv.wait(rg, [request] { return request->state != XTBLOOM_REQUEST_PENDING; });
  }
  delete request;
}

// Admits one submission and queues it for the context's worker (reference
// xtbloom.h:892-925): admission errors return directly and leave the request
// and every output untouched; an admitted call leaves the request PENDING
// and returns at once.  Host inputs are staged during admission, so the
// caller may reuse them on return; output buffers (and device inputs, staged
// here as well) stay borrowed until COMPLETE.  The caller's result
// descriptor is copied and never modified: completed result flags go to the
// request.
static xtbloom_status_t enqueue_impl(xtbloom_context_t* context, xtbloom_plan_t* plan,
                                     const xtbloom_batch_t* batch, const xtbloom_compute_options_t* options,
                                     const xtbloom_batch_result_t* result, xtbloom_request_t* request) {
  if (request == nullptr) return fail(XTBLOOM_STATUS_INVALID_ARGUMENT, "request is NULL");
  if (context == nullptr || context->device == nullptr)
    return fail(XTBLOOM_STATUS_INVALID_ARGUMENT, "context is NULL");
  if (request->context != context)
    return fail(XTBLOOM_STATUS_INVALID_ARGUMENT, "request belongs to another context");
  if (!valid_header(result, XTBLOOM_BATCH_RESULT_V1_SIZE))
    return fail(XTBLOOM_STATUS_INVALID_ARGUMENT, "batch result is NULL, too small, or has a different api_version");
  std::unique_lock<std::mutex> guard(context->lock);
  {
    std::lock_guard<std::mutex> rg(request->m);
    if (request->state == XTBLOOM_REQUEST_PENDING)
      return fail(XTBLOOM_STATUS_INVALID_ARGUMENT, "request already has a pending submission");
  }
  std::unique_ptr<Job> job;
  try {
    job.reset(new Job());
  } catch (const std::bad_alloc&) {
    return fail(XTBLOOM_STATUS_ALLOCATION_FAILED, "host allocation failed");
  }
  const xtbloom_status_t s = admit(context, plan, batch, options, result, nullptr, false, *job);
  if (s != XTBLOOM_STATUS_SUCCESS) return s;
  job->request = request;
  try {
    if (!context->worker.joinable()) context->worker = std::thread(worker_loop, context);
    context->queue.push_back(job.get());
  } catch (...) {
    return fail(XTBLOOM_STATUS_INTERNAL_ERROR, "the submission worker could not be started");
  }
  job.release();
  {
    std::lock_guard<std::mutex> rg(request->m);
    request->state = XTBLOOM_REQUEST_PENDING;
    request->completion_status = XTBLOOM_STATUS_SUCCESS;
    request->result_flags = 0;
    request->error.clear();
  }
  context->cv.notify_all();
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
  std::unique_lock<std::mutex> guard(context->lock);
  drain(context, guard);  // the reservation runs on the device
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

// queued submissions may still reference the plan: they complete first
void xtbloom_plan_destroy(xtbloom_plan_t* plan) {
  if (plan == nullptr) return;
  {
    std::unique_lock<std::mutex> guard(plan->context->lock);
    drain(plan->context, guard);
  }
  delete plan;
}

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
  if (s == XTBLOOM_STATUS_SUCCESS) {
    o->memory_space = XTBLOOM_MEMORY_HOST;
    o->device_id = -1;
  }
  return s;
}

xtbloom_status_t xtbloom_result_owner_create(const xtbloom_result_owner_options_t* options,
                                             xtbloom_result_owner_t** owner) {
  if (owner == nullptr) return fail(XTBLOOM_STATUS_INVALID_ARGUMENT, "result owner output is NULL");
  *owner = nullptr;
  if (!valid_header(options, XTBLOOM_RESULT_OWNER_OPTIONS_V1_SIZE) || options->reserved != 0)
    return fail(XTBLOOM_STATUS_INVALID_ARGUMENT,
                "result owner options are NULL, too small, have a different api_version or a nonzero reserved field");
  if (options->size_bytes == 0) return fail(XTBLOOM_STATUS_INVALID_ARGUMENT, "result arena size must be nonzero");
  const bool host = options->memory_space == XTBLOOM_MEMORY_HOST;
  if (host && options->device_id != -1)
    return fail(XTBLOOM_STATUS_INVALID_ARGUMENT, "a host result arena takes device_id -1");
  if (!host && options->memory_space != XTBLOOM_MEMORY_ROCM_DEVICE)
    return fail(XTBLOOM_STATUS_INVALID_ARGUMENT, "result arenas live in host or ROCm device memory");
  if (!host && options->device_id < 0)
    return fail(XTBLOOM_STATUS_INVALID_ARGUMENT, "a device result arena needs a device ordinal");
  std::unique_ptr<xtbloom_result_owner_t> o;
  try {
    o.reset(new xtbloom_result_owner_t());
  } catch (const std::bad_alloc&) {
    return fail(XTBLOOM_STATUS_ALLOCATION_FAILED, "result owner allocation failed");
  }
  o->memory_space = options->memory_space;
  o->device_id = options->device_id;
  o->size = options->size_bytes;
  if (host) {
    const uint64_t rounded = (o->size + kArenaAlign - 1) / kArenaAlign * kArenaAlign;
    o->data = std::aligned_alloc(kArenaAlign, rounded);
    if (o->data == nullptr) return fail(XTBLOOM_STATUS_ALLOCATION_FAILED, "host result arena allocation failed");
    std::memset(o->data, 0, rounded);
  } else {
    DeviceScope scope(o->device_id);
    if (!scope.ok) return fail(XTBLOOM_STATUS_INVALID_ARGUMENT, "device result arena: no such device");
    if (hipMalloc(&o->data, o->size) != hipSuccess || o->data == nullptr) {
      o->data = nullptr;
      return fail(XTBLOOM_STATUS_ALLOCATION_FAILED, "device result arena allocation failed");
    }
    if (hipMemset(o->data, 0, o->size) != hipSuccess) {
      (void)hipFree(o->data);
      o->data = nullptr;
      return fail(XTBLOOM_STATUS_INTERNAL_ERROR, "device result arena initialization failed");
    }
  }
  *owner = o.release();
  g_last_error.clear();
  return XTBLOOM_STATUS_SUCCESS;
}

xtbloom_status_t xtbloom_result_owner_buffer(const xtbloom_result_owner_t* owner, xtbloom_buffer_t* buffer) {
  if (owner == nullptr || buffer == nullptr)
    return fail(XTBLOOM_STATUS_INVALID_ARGUMENT, "result owner or buffer is NULL");
  buffer->data = owner->data;
  buffer->size_bytes = (size_t)owner->size;
  buffer->memory_space = owner->memory_space;
  buffer->reserved = 0;
  g_last_error.clear();
  return XTBLOOM_STATUS_SUCCESS;
}

void xtbloom_result_owner_retain(xtbloom_result_owner_t* owner) {
  if (owner != nullptr) owner->refs.fetch_add(1, std::memory_order_relaxed);
}

void xtbloom_result_owner_release(xtbloom_result_owner_t* owner) {
  if (owner == nullptr) return;
  if (owner->refs.fetch_sub(1, std::memory_order_acq_rel) == 1) delete owner;
}

xtbloom_status_t xtbloom_result_owner_export_dltensor(const xtbloom_result_owner_t* owner,
                                                      const xtbloom_dlpack_view_t* view, int version,
                                                      void** out_managed) {
  if (out_managed == nullptr) return fail(XTBLOOM_STATUS_INVALID_ARGUMENT, "managed tensor output is NULL");
  *out_managed = nullptr;
  if (owner == nullptr) return fail(XTBLOOM_STATUS_INVALID_ARGUMENT, "result owner is NULL");
  if (version != 0 && version != 1)
    return fail(XTBLOOM_STATUS_INVALID_ARGUMENT, "DLPack export version must be 0 (legacy) or 1 (versioned)");
  if (!valid_header(view, XTBLOOM_DLPACK_VIEW_V1_SIZE) || view->reserved != 0)
    return fail(XTBLOOM_STATUS_INVALID_ARGUMENT,
                "DLPack view is NULL, too small, has a different api_version or a nonzero reserved field");
  // scalar int8/16/32/64, uint8, float32/64 (reference xtbloom.h:1067-1076)
  const int32_t code = view->dtype_code, bits = view->dtype_bits;
  const bool dtype_ok = view->dtype_lanes == 1 &&
                        ((code == kDLInt && (bits == 8 || bits == 16 || bits == 32 || bits == 64)) ||
                         (code == kDLUInt && bits == 8) || (code == kDLFloat && (bits == 32 || bits == 64)));
  if (!dtype_ok) return fail(XTBLOOM_STATUS_INVALID_ARGUMENT, "unsupported DLPack dtype");
  if (view->ndim < 0 || view->ndim > XTBLOOM_DLPACK_MAX_NDIM)
    return fail(XTBLOOM_STATUS_INVALID_ARGUMENT, "DLPack ndim must be 0..8");
  if (view->ndim > 0 && view->shape == nullptr) return fail(XTBLOOM_STATUS_INVALID_ARGUMENT, "DLPack shape is NULL");
  const uint64_t item = (uint64_t)bits / 8;
  if (view->byte_offset % item != 0)
    return fail(XTBLOOM_STATUS_INVALID_ARGUMENT, "DLPack byte offset breaks the dtype alignment");
  uint64_t count = 1;
  bool empty = false;
  for (int32_t d = 0; d < view->ndim; ++d) {
    const int64_t extent = view->shape[d];
    if (extent < 0) return fail(XTBLOOM_STATUS_INVALID_ARGUMENT, "negative DLPack extent");
    if (extent == 0) empty = true;
  }
  if (!empty) {
    for (int32_t d = 0; d < view->ndim; ++d) {
      const uint64_t extent = (uint64_t)view->shape[d];
      if (count > std::numeric_limits<uint64_t>::max() / extent)
        return fail(XTBLOOM_STATUS_INVALID_ARGUMENT, "DLPack extent product overflows");
      count *= extent;
    }
    if (count > (std::numeric_limits<uint64_t>::max() - view->byte_offset) / item ||
        view->byte_offset + count * item > owner->size)
      return fail(XTBLOOM_STATUS_INVALID_ARGUMENT, "DLPack view exceeds the result arena");
  } else if (view->byte_offset > owner->size) {
    return fail(XTBLOOM_STATUS_INVALID_ARGUMENT, "DLPack view exceeds the result arena");
  }
  ManagedExport* m = nullptr;
  try {
    m = new ManagedExport();
  } catch (const std::bad_alloc&) {
    return fail(XTBLOOM_STATUS_ALLOCATION_FAILED, "managed tensor allocation failed");
  }
  std::memset(&m->legacy, 0, sizeof(m->legacy));
  std::memset(&m->versioned, 0, sizeof(m->versioned));
  for (int32_t d = 0; d < view->ndim; ++d) m->shape[d] = view->shape[d];
  m->owner = const_cast<xtbloom_result_owner_t*>(owner);
  DLTensorLayout t;
  std::memset(&t, 0, sizeof(t));
  t.data = empty ? nullptr : static_cast<char*>(owner->data) + view->byte_offset;
  const bool host = owner->memory_space == XTBLOOM_MEMORY_HOST;
  t.device.device_type = host ? kDLCPU : kDLROCM;
  t.device.device_id = host ? 0 : owner->device_id;
  t.ndim = view->ndim;
  t.dtype.code = (uint8_t)code;
  t.dtype.bits = (uint8_t)bits;
  t.dtype.lanes = 1;
  t.shape = view->ndim > 0 ? m->shape : nullptr;
  t.strides = nullptr;
  t.byte_offset = 0;
  xtbloom_result_owner_retain(m->owner);
  if (version == 1) {
    m->versioned.version_major = 1;
    m->versioned.version_minor = 0;
    m->versioned.manager_ctx = m;
    m->versioned.deleter = &ManagedExport::delete_versioned;
    m->versioned.flags = 0;
    m->versioned.dl_tensor = t;
    *out_managed = &m->versioned;
  } else {
    m->legacy.dl_tensor = t;
    m->legacy.manager_ctx = m;
    m->legacy.deleter = &ManagedExport::delete_legacy;
    *out_managed = &m->legacy;
  }
  g_last_error.clear();
  return XTBLOOM_STATUS_SUCCESS;
}

}  // extern "C"
