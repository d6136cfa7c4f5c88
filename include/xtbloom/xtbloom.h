/*
 * xtbloom_amd -- the caller-buffer C ABI of xTBloom, served by an MI355X (gfx950) backend.
 *
 * This header restates the reference ABI (include/xtbloom/xtbloom.h of the
 * reference) so that an existing caller can load libxtbloom_amd.so in place of
 * libxtbloom.so: every struct keeps the reference's field order, widths and
 * byte offsets (asserted below for LP64), every tag keeps its value, and every
 * entry point keeps its name, argument meaning and status semantics.  The
 * backend implemented here is XTBLOOM_BACKEND_ROCM (reserved in the reference
 * at xtbloom.h:96-97); AUTO resolves to ROCm.
 *
 * Plain pointers and sizes only; no torch or HIP types appear in signatures
 * (the stream is a hipStream_t cast to void*, as in the reference).
 */
#ifndef XTBLOOM_XTBLOOM_H
#define XTBLOOM_XTBLOOM_H

#include <stddef.h>
#include <stdint.h>

#if defined(__GNUC__) || defined(__clang__)
#define XTBLOOM_API __attribute__((visibility("default")))
#else
#define XTBLOOM_API
#endif

#ifdef __cplusplus
extern "C" {
#endif

/* Version of this library (the reference generates this from its packaging metadata). */
#define XTBLOOM_VERSION_MAJOR 0
#define XTBLOOM_VERSION_MINOR 1
#define XTBLOOM_VERSION_PATCH 0
#define XTBLOOM_VERSION_STRING "0.1.0+rocm.gfx950"

#define XTBLOOM_API_VERSION 1u
#define XTBLOOM_KELVIN_TO_HARTREE 3.166808578545117e-6
#define XTBLOOM_DEFAULT_ELECTRONIC_TEMPERATURE (300.0 * XTBLOOM_KELVIN_TO_HARTREE)

typedef struct xtbloom_context xtbloom_context_t;
typedef struct xtbloom_plan xtbloom_plan_t;
typedef struct xtbloom_request xtbloom_request_t;
typedef struct xtbloom_result_owner xtbloom_result_owner_t;

typedef int32_t xtbloom_status_t; /* reference xtbloom.h:59-71 */
enum xtbloom_status_value {
  XTBLOOM_STATUS_SUCCESS = 0,
  XTBLOOM_STATUS_INVALID_ARGUMENT = 1,
  XTBLOOM_STATUS_BACKEND_UNAVAILABLE = 2,
  XTBLOOM_STATUS_NOT_SUPPORTED = 3,
  XTBLOOM_STATUS_ALLOCATION_FAILED = 4,
  XTBLOOM_STATUS_NOT_IMPLEMENTED = 5,
  XTBLOOM_STATUS_INTERNAL_ERROR = 6,
  XTBLOOM_STATUS_SCC_NOT_CONVERGED = 7,
  XTBLOOM_STATUS_EIGENSOLVER_FAILED = 8
};

typedef int32_t xtbloom_request_state_t;
enum xtbloom_request_state_value {
  XTBLOOM_REQUEST_IDLE = 0,
  XTBLOOM_REQUEST_PENDING = 1,
  XTBLOOM_REQUEST_COMPLETE = 2
};

typedef int32_t xtbloom_backend_t; /* reference xtbloom.h:90-98 */
enum xtbloom_backend_value {
  XTBLOOM_BACKEND_AUTO = 0,
  XTBLOOM_BACKEND_CPU = 1,
  XTBLOOM_BACKEND_CUDA = 2,
  XTBLOOM_BACKEND_ROCM = 3
};

typedef int32_t xtbloom_memory_space_t; /* reference xtbloom.h:100-105 */
enum xtbloom_memory_space_value {
  XTBLOOM_MEMORY_HOST = 0,
  XTBLOOM_MEMORY_CUDA_DEVICE = 1,
  XTBLOOM_MEMORY_ROCM_DEVICE = 2
};

typedef int32_t xtbloom_model_t;
enum xtbloom_model_value { XTBLOOM_MODEL_GFN1_XTB = 1, XTBLOOM_MODEL_GFN2_XTB = 2 };

typedef int32_t xtbloom_scc_start_mode_t;
enum xtbloom_scc_start_mode_value { XTBLOOM_SCC_START_FRESH = 1, XTBLOOM_SCC_START_WARM = 2 };

typedef int32_t xtbloom_scc_mixer_t;
enum xtbloom_scc_mixer_value { XTBLOOM_SCC_MIXER_MODIFIED_BROYDEN = 1 };

typedef int32_t xtbloom_determinism_t;
enum xtbloom_determinism_value {
  XTBLOOM_DETERMINISM_DEFAULT = 0,
  XTBLOOM_DETERMINISM_REPRODUCIBLE = 1
};

typedef int32_t xtbloom_compute_flag_t; /* reference xtbloom.h:134-142 */
enum xtbloom_compute_flag_value {
  XTBLOOM_COMPUTE_ENERGY = 1 << 0,
  XTBLOOM_COMPUTE_FORCES = 1 << 1,
  XTBLOOM_COMPUTE_ATOMIC_CHARGES = 1 << 2,
  XTBLOOM_COMPUTE_POINT_CHARGE_FORCES = 1 << 3,
  XTBLOOM_COMPUTE_DIPOLE_MOMENTS = 1 << 4
};

typedef int32_t xtbloom_result_flag_t;
enum xtbloom_result_flag_value {
  XTBLOOM_RESULT_FORCES_EXCLUDE_EXTERNAL_OPERATOR_DERIVATIVES = 1 << 0,
  XTBLOOM_RESULT_DIPOLE_MOMENTS = 1 << 4
};

typedef int32_t xtbloom_interaction_type_t;
enum xtbloom_interaction_type_value {
  XTBLOOM_INTERACTION_NONE = 0,
  XTBLOOM_INTERACTION_ELECTRIC_FIELD = 0x0101,
  XTBLOOM_INTERACTION_ELECTRIC_FIELD_GRADIENT = 0x0102,
  XTBLOOM_INTERACTION_POINT_CHARGES_MULTIPOLE = 0x0103,
  XTBLOOM_INTERACTION_ATOMIC_POTENTIAL_GRID = 0x0104,
  XTBLOOM_INTERACTION_ALPB_SOLVATION = 0x0201,
  XTBLOOM_INTERACTION_GBSA_SOLVATION = 0x0202,
  XTBLOOM_INTERACTION_GB_SOLVATION = 0x0203,
  XTBLOOM_INTERACTION_GBE_SOLVATION = 0x0204,
  XTBLOOM_INTERACTION_DDX_SOLVATION = 0x0205,
  XTBLOOM_INTERACTION_D3_DISPERSION = 0x0301,
  XTBLOOM_INTERACTION_D4_VARIANT_DISPERSION = 0x0302,
  XTBLOOM_INTERACTION_HALOGEN_BOND = 0x0401
};

typedef int32_t xtbloom_periodic_axes_t;
enum xtbloom_periodic_axes_value {
  XTBLOOM_PERIODIC_AXES_NONE = 0,
  XTBLOOM_PERIODIC_AXIS_X = 1 << 0,
  XTBLOOM_PERIODIC_AXIS_Y = 1 << 1,
  XTBLOOM_PERIODIC_AXIS_Z = 1 << 2,
  XTBLOOM_PERIODIC_AXES_XYZ = 7
};

typedef struct xtbloom_interaction { /* reference xtbloom.h:196-205 */
  xtbloom_interaction_type_t type;
  uint32_t flags;
  int64_t system_index;
  uint64_t payload_offset;
  uint64_t payload_size;
} xtbloom_interaction_t;

typedef struct xtbloom_context_options { /* reference xtbloom.h:276-289 */
  uint32_t struct_size;
  uint32_t api_version;
  xtbloom_backend_t backend;
  int32_t device_id;
  int32_t cpu_threads;
  uint32_t reserved;
  void* stream; /* hipStream_t cast to void*; NULL selects the null stream */
} xtbloom_context_options_t;
#define XTBLOOM_CONTEXT_OPTIONS_V1_SIZE \
  (offsetof(xtbloom_context_options_t, stream) + sizeof(void*))

typedef struct xtbloom_const_buffer { /* reference xtbloom.h:294-299 */
  const void* data;
  size_t size_bytes;
  xtbloom_memory_space_t memory_space;
  uint32_t reserved;
} xtbloom_const_buffer_t;

typedef struct xtbloom_buffer { /* reference xtbloom.h:302-307 */
  void* data;
  size_t size_bytes;
  xtbloom_memory_space_t memory_space;
  uint32_t reserved;
} xtbloom_buffer_t;

/* Ragged molecular batch (reference xtbloom.h:437-487). */
typedef struct xtbloom_batch {
  uint32_t struct_size;
  uint32_t api_version;
  int64_t batch_size;
  int64_t total_atoms;
  int64_t total_point_charges;
  int64_t total_charge_response_elements;
  xtbloom_const_buffer_t atom_offsets;       /* int64 [B+1] */
  xtbloom_const_buffer_t atomic_numbers;     /* int32 [N] */
  xtbloom_const_buffer_t positions;          /* float64 [3N], bohr */
  xtbloom_const_buffer_t molecular_charges;  /* float64 [B] */
  xtbloom_const_buffer_t unpaired_electrons; /* int32 [B] */
  xtbloom_const_buffer_t point_charge_offsets;   /* int64 [B+1] */
  xtbloom_const_buffer_t point_charge_positions; /* float64 [3P] */
  xtbloom_const_buffer_t point_charge_values;    /* float64 [P] */
  xtbloom_const_buffer_t point_charge_gammas;    /* float64 [P] */
  xtbloom_const_buffer_t atomic_potential_shifts;
  xtbloom_const_buffer_t charge_response_offsets;
  xtbloom_const_buffer_t charge_response_matrix;
  xtbloom_const_buffer_t spin_channels;          /* ABI v2 */
  int64_t total_interactions;                    /* ABI v3 */
  xtbloom_const_buffer_t interaction_descriptors;
  xtbloom_const_buffer_t interaction_payload;
  xtbloom_const_buffer_t cell_matrices;          /* ABI v4 */
  xtbloom_const_buffer_t periodic_axes;
} xtbloom_batch_t;
#define XTBLOOM_BATCH_V1_SIZE \
  (offsetof(xtbloom_batch_t, charge_response_matrix) + sizeof(xtbloom_const_buffer_t))
#define XTBLOOM_BATCH_V2_SIZE \
  (offsetof(xtbloom_batch_t, spin_channels) + sizeof(xtbloom_const_buffer_t))
#define XTBLOOM_BATCH_V3_SIZE \
  (offsetof(xtbloom_batch_t, interaction_payload) + sizeof(xtbloom_const_buffer_t))
#define XTBLOOM_BATCH_V4_SIZE \
  (offsetof(xtbloom_batch_t, periodic_axes) + sizeof(xtbloom_const_buffer_t))

/* SCC policy (reference xtbloom.h:525-552; defaults reference api.cpp:150-173). */
typedef struct xtbloom_compute_options {
  uint32_t struct_size;
  uint32_t api_version;
  xtbloom_model_t model;
  uint32_t flags;
  int32_t max_scc_iterations;
  uint32_t reserved;
  double charge_tolerance;
  double energy_tolerance;
  double electronic_temperature; /* k_B*T in Hartree */
  xtbloom_scc_start_mode_t scc_start_mode; /* ABI v2 */
  uint32_t reserved_v2;
  xtbloom_scc_mixer_t scc_mixer; /* ABI v3 */
  int32_t scc_mixer_history;
  double scc_mixer_damping;
  xtbloom_determinism_t determinism;
  uint32_t reserved_v3;
} xtbloom_compute_options_t;
#define XTBLOOM_COMPUTE_OPTIONS_V1_SIZE \
  (offsetof(xtbloom_compute_options_t, electronic_temperature) + sizeof(double))
#define XTBLOOM_COMPUTE_OPTIONS_V2_SIZE \
  (offsetof(xtbloom_compute_options_t, reserved_v2) + sizeof(uint32_t))
#define XTBLOOM_COMPUTE_OPTIONS_V3_SIZE \
  (offsetof(xtbloom_compute_options_t, reserved_v3) + sizeof(uint32_t))

/* Caller-allocated results (reference xtbloom.h:664-687). */
typedef struct xtbloom_batch_result {
  uint32_t struct_size;
  uint32_t api_version;
  uint32_t flags;
  uint32_t reserved;
  xtbloom_buffer_t energies;            /* float64 [B] */
  xtbloom_buffer_t forces;              /* float64 [3N] */
  xtbloom_buffer_t atomic_charges;      /* float64 [N] */
  xtbloom_buffer_t point_charge_forces; /* float64 [3P] */
  xtbloom_buffer_t scc_iterations;      /* int32 [B] */
  xtbloom_buffer_t scc_converged;       /* uint8 [B] */
  xtbloom_buffer_t per_system_status;   /* int32 [B] */
  xtbloom_buffer_t dipole_moments;      /* ABI v2: float64 [3B] */
  xtbloom_buffer_t quadrupole_moments;
  xtbloom_buffer_t wiberg_orders;
  xtbloom_buffer_t spin_populations;
} xtbloom_batch_result_t;
#define XTBLOOM_BATCH_RESULT_V1_SIZE \
  (offsetof(xtbloom_batch_result_t, per_system_status) + sizeof(xtbloom_buffer_t))
#define XTBLOOM_BATCH_RESULT_V2_SIZE \
  (offsetof(xtbloom_batch_result_t, spin_populations) + sizeof(xtbloom_buffer_t))

typedef struct xtbloom_workspace_query { /* reference xtbloom.h:739-749 */
  uint32_t struct_size;
  uint32_t api_version;
  uint32_t compute_flags;
  uint32_t reserved;
  uint64_t host_required_bytes;
  uint32_t host_required_alignment;
  uint64_t device_required_bytes;
  uint32_t device_required_alignment;
  uint32_t reserved_v2;
} xtbloom_workspace_query_t;
#define XTBLOOM_WORKSPACE_QUERY_V1_SIZE \
  (offsetof(xtbloom_workspace_query_t, reserved_v2) + sizeof(uint32_t))

typedef struct xtbloom_request_info { /* reference xtbloom.h:775-782 */
  uint32_t struct_size;
  uint32_t api_version;
  xtbloom_request_state_t state;
  xtbloom_status_t completion_status;
  uint32_t result_flags;
  uint32_t reserved;
} xtbloom_request_info_t;
#define XTBLOOM_REQUEST_INFO_V1_SIZE (offsetof(xtbloom_request_info_t, reserved) + sizeof(uint32_t))

typedef struct xtbloom_result_owner_options { /* reference xtbloom.h:1047-1057 */
  uint32_t struct_size;
  uint32_t api_version;
  xtbloom_memory_space_t memory_space;
  int32_t device_id;
  uint64_t size_bytes;
  uint32_t reserved;
} xtbloom_result_owner_options_t;
#define XTBLOOM_RESULT_OWNER_OPTIONS_V1_SIZE \
  (offsetof(xtbloom_result_owner_options_t, reserved) + sizeof(uint32_t))

typedef struct xtbloom_dlpack_view { /* reference xtbloom.h:1074-1085 */
  uint32_t struct_size;
  uint32_t api_version;
  uint64_t byte_offset;
  int32_t dtype_code;
  int32_t dtype_bits;
  int32_t dtype_lanes;
  int32_t ndim;
  uint32_t reserved;
  const int64_t* shape;
} xtbloom_dlpack_view_t;
#define XTBLOOM_DLPACK_MAX_NDIM 8
#define XTBLOOM_DLPACK_VIEW_V1_SIZE (offsetof(xtbloom_dlpack_view_t, shape) + sizeof(const int64_t*))

#if UINTPTR_MAX == UINT64_MAX
#if defined(__cplusplus)
#define XTBLOOM_LAYOUT_CHECK(c, m) static_assert(c, m)
#else
#define XTBLOOM_LAYOUT_CHECK(c, m) _Static_assert(c, m)
#endif
/* Byte images asserted by the reference for LP64 (reference xtbloom.h:320-343). */
XTBLOOM_LAYOUT_CHECK(XTBLOOM_CONTEXT_OPTIONS_V1_SIZE == 32u, "context options image");
XTBLOOM_LAYOUT_CHECK(sizeof(xtbloom_buffer_t) == 24u, "buffer image");
XTBLOOM_LAYOUT_CHECK(XTBLOOM_BATCH_V1_SIZE == 328u, "batch v1 image");
XTBLOOM_LAYOUT_CHECK(XTBLOOM_BATCH_V2_SIZE == 352u, "batch v2 image");
XTBLOOM_LAYOUT_CHECK(XTBLOOM_BATCH_V3_SIZE == 408u, "batch v3 image");
XTBLOOM_LAYOUT_CHECK(XTBLOOM_BATCH_V4_SIZE == 456u, "batch v4 image");
XTBLOOM_LAYOUT_CHECK(XTBLOOM_COMPUTE_OPTIONS_V1_SIZE == 48u, "options v1 image");
XTBLOOM_LAYOUT_CHECK(XTBLOOM_COMPUTE_OPTIONS_V2_SIZE == 56u, "options v2 image");
XTBLOOM_LAYOUT_CHECK(XTBLOOM_COMPUTE_OPTIONS_V3_SIZE == 80u, "options v3 image");
XTBLOOM_LAYOUT_CHECK(XTBLOOM_BATCH_RESULT_V1_SIZE == 184u, "result v1 image");
XTBLOOM_LAYOUT_CHECK(XTBLOOM_BATCH_RESULT_V2_SIZE == 280u, "result v2 image");
XTBLOOM_LAYOUT_CHECK(sizeof(xtbloom_interaction_t) == 32u, "interaction image");
XTBLOOM_LAYOUT_CHECK(XTBLOOM_DLPACK_VIEW_V1_SIZE == 48u, "dlpack view image");
#undef XTBLOOM_LAYOUT_CHECK
#endif

/* --- library metadata and diagnostics (reference xtbloom.h:821-826, api.cpp:45-132) --- */
XTBLOOM_API const char* xtbloom_version_string(void);
XTBLOOM_API const char* xtbloom_status_string(xtbloom_status_t status);
/* Thread-local message of the last failing call on this thread. */
XTBLOOM_API const char* xtbloom_get_last_error(void);

/* --- struct initializers (reference xtbloom.h:827-837, defaults api.cpp:140-200) --- */
XTBLOOM_API xtbloom_status_t xtbloom_context_options_init(xtbloom_context_options_t* options,
                                                          size_t struct_size);
XTBLOOM_API xtbloom_status_t xtbloom_batch_init(xtbloom_batch_t* batch, size_t struct_size);
XTBLOOM_API xtbloom_status_t xtbloom_compute_options_init(xtbloom_compute_options_t* options,
                                                          size_t struct_size);
XTBLOOM_API xtbloom_status_t xtbloom_batch_result_init(xtbloom_batch_result_t* result,
                                                       size_t struct_size);
XTBLOOM_API xtbloom_status_t xtbloom_workspace_query_init(xtbloom_workspace_query_t* query,
                                                          size_t struct_size);
XTBLOOM_API xtbloom_status_t xtbloom_request_info_init(xtbloom_request_info_t* info,
                                                       size_t struct_size);

/* --- contexts (reference xtbloom.h:839-846; backend.cpp:31-115) --- */
/* ROCM and AUTO create an MI355X context; CPU and CUDA return BACKEND_UNAVAILABLE. */
XTBLOOM_API xtbloom_status_t xtbloom_context_create(const xtbloom_context_options_t* options,
                                                    xtbloom_context_t** context);
XTBLOOM_API void xtbloom_context_destroy(xtbloom_context_t* context);
XTBLOOM_API xtbloom_backend_t xtbloom_context_get_backend(const xtbloom_context_t* context);
XTBLOOM_API int32_t xtbloom_context_get_device_id(const xtbloom_context_t* context);

/*
 * Synchronous batched GFN2 single point (reference xtbloom.h:896-899, api.cpp:486).
 * Host and ROCm-device buffers are accepted.  Validation happens before any
 * output is written; per-system SCC/eigensolver failures are data (SUCCESS
 * return, NaN slices, per_system_status).
 */
XTBLOOM_API xtbloom_status_t xtbloom_compute(xtbloom_context_t* context,
                                             const xtbloom_batch_t* batch,
                                             const xtbloom_compute_options_t* options,
                                             xtbloom_batch_result_t* result);

/* --- asynchronous requests (reference xtbloom.h:850-878, 892-937).  An
 * admitted enqueue leaves the request PENDING and returns at once; the
 * context's worker thread runs submissions in order and completes the request
 * (status, error text and result flags in the request).  Host inputs are
 * staged at admission; output buffers stay borrowed until COMPLETE.
 * Admission errors return directly and leave the request unchanged. --- */
XTBLOOM_API xtbloom_status_t xtbloom_request_create(xtbloom_context_t* context,
                                                    xtbloom_request_t** request);
XTBLOOM_API xtbloom_status_t xtbloom_request_query(xtbloom_request_t* request,
                                                   xtbloom_request_info_t* info);
XTBLOOM_API xtbloom_status_t xtbloom_request_wait(xtbloom_request_t* request,
                                                  xtbloom_request_info_t* info);
XTBLOOM_API const char* xtbloom_request_get_error(const xtbloom_request_t* request);
XTBLOOM_API void xtbloom_request_destroy(xtbloom_request_t* request);
XTBLOOM_API xtbloom_status_t xtbloom_compute_enqueue(xtbloom_context_t* context,
                                                     const xtbloom_batch_t* batch,
                                                     const xtbloom_compute_options_t* options,
                                                     const xtbloom_batch_result_t* result,
                                                     xtbloom_request_t* request);

/* --- fixed-topology plans (reference xtbloom.h:960-1009): the plan binds the
 * topology and compute policy, keeps the packed host topology and its own
 * strict-WARM checkpoint, and reserves the device workspace at creation. --- */
XTBLOOM_API xtbloom_status_t xtbloom_plan_create(xtbloom_context_t* context,
                                                 const xtbloom_batch_t* batch,
                                                 const xtbloom_compute_options_t* options,
                                                 xtbloom_plan_t** plan);
XTBLOOM_API void xtbloom_plan_destroy(xtbloom_plan_t* plan);
XTBLOOM_API xtbloom_status_t xtbloom_plan_query_workspace(const xtbloom_plan_t* plan,
                                                          xtbloom_workspace_query_t* query);
XTBLOOM_API xtbloom_status_t xtbloom_plan_compute(xtbloom_plan_t* plan,
                                                  const xtbloom_batch_t* batch,
                                                  const xtbloom_compute_options_t* options,
                                                  xtbloom_batch_result_t* result);
XTBLOOM_API xtbloom_status_t xtbloom_plan_compute_enqueue(xtbloom_plan_t* plan,
                                                          const xtbloom_batch_t* batch,
                                                          const xtbloom_compute_options_t* options,
                                                          const xtbloom_batch_result_t* result,
                                                          xtbloom_request_t* request);

/* --- result owners (reference xtbloom.h:1011-1145): reference-counted
 * arenas in host or ROCm device memory, exported as DLPack 1.0 tensors. --- */
XTBLOOM_API xtbloom_status_t
xtbloom_result_owner_options_init(xtbloom_result_owner_options_t* options, size_t struct_size);
XTBLOOM_API xtbloom_status_t xtbloom_result_owner_create(
    const xtbloom_result_owner_options_t* options, xtbloom_result_owner_t** owner);
XTBLOOM_API xtbloom_status_t xtbloom_result_owner_buffer(const xtbloom_result_owner_t* owner,
                                                         xtbloom_buffer_t* buffer);
XTBLOOM_API void xtbloom_result_owner_retain(xtbloom_result_owner_t* owner);
XTBLOOM_API void xtbloom_result_owner_release(xtbloom_result_owner_t* owner);
XTBLOOM_API xtbloom_status_t xtbloom_result_owner_export_dltensor(
    const xtbloom_result_owner_t* owner, const xtbloom_dlpack_view_t* view, int version,
    void** out_managed);

#ifdef __cplusplus
}
#endif

#endif
