"""ctypes mirror of the xtbloom C ABI for the MI355X library (``lib/libxtbloom_amd.so``).

Struct images are byte-identical to the reference ABI (reference
python/xtbloom/library.py:128-240 binds the same layouts); the library path can
be overridden with ``XTBLOOM_AMD_LIBRARY``.  Loading fails loudly when the
compiled HIP library is missing -- there is no fallback implementation.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_LIBRARY = os.path.join(HERE, "lib", "libxtbloom_amd.so")

API_VERSION = 1
KELVIN_TO_HARTREE = 3.166808578545117e-6

STATUS_SUCCESS = 0
STATUS_INVALID_ARGUMENT = 1
STATUS_BACKEND_UNAVAILABLE = 2
STATUS_NOT_SUPPORTED = 3
STATUS_ALLOCATION_FAILED = 4
STATUS_NOT_IMPLEMENTED = 5
STATUS_INTERNAL_ERROR = 6
STATUS_SCC_NOT_CONVERGED = 7
STATUS_EIGENSOLVER_FAILED = 8

NO_ALLOC_WORKSPACE_BYTES = 64 << 20  # XTBLOOM_AMD_NO_ALLOC_WORKSPACE_BYTES

BACKEND_AUTO = 0
BACKEND_CPU = 1
BACKEND_CUDA = 2
BACKEND_ROCM = 3

MEMORY_HOST = 0
MEMORY_CUDA_DEVICE = 1
MEMORY_ROCM_DEVICE = 2

MODEL_GFN1_XTB = 1
MODEL_GFN2_XTB = 2

COMPUTE_ENERGY = 1 << 0
COMPUTE_FORCES = 1 << 1
COMPUTE_ATOMIC_CHARGES = 1 << 2
COMPUTE_POINT_CHARGE_FORCES = 1 << 3
COMPUTE_DIPOLE_MOMENTS = 1 << 4

# xtbloom_batch_result_t.flags
RESULT_FORCES_EXCLUDE_EXTERNAL_OPERATOR_DERIVATIVES = 1 << 0
RESULT_DIPOLE_MOMENTS = 1 << 4

SCC_START_FRESH = 1
SCC_START_WARM = 2
SCC_MIXER_MODIFIED_BROYDEN = 1
DETERMINISM_DEFAULT = 0
DETERMINISM_REPRODUCIBLE = 1
DEFAULT_SCC_MIXER_HISTORY = 8
DEFAULT_SCC_MIXER_DAMPING = 0.4

# interaction attachment tags (ABI-v3 batch suffix); the uniform electric field
# (block_version 1: int32 version, int32 reserved, three doubles) is executed,
# the others are reserved
INTERACTION_NONE = 0
INTERACTION_ELECTRIC_FIELD = 0x0101
INTERACTION_ELECTRIC_FIELD_GRADIENT = 0x0102
INTERACTION_ALPB_SOLVATION = 0x0201
ELECTRIC_FIELD_BLOCK_BYTES = 32


class ContextOptions(ctypes.Structure):
    _fields_ = [("struct_size", ctypes.c_uint32), ("api_version", ctypes.c_uint32),
                ("backend", ctypes.c_int32), ("device_id", ctypes.c_int32),
                ("cpu_threads", ctypes.c_int32), ("reserved", ctypes.c_uint32),
                ("stream", ctypes.c_void_p)]


class ConstBuffer(ctypes.Structure):
    _fields_ = [("data", ctypes.c_void_p), ("size_bytes", ctypes.c_size_t),
                ("memory_space", ctypes.c_int32), ("reserved", ctypes.c_uint32)]


class Buffer(ctypes.Structure):
    _fields_ = [("data", ctypes.c_void_p), ("size_bytes", ctypes.c_size_t),
                ("memory_space", ctypes.c_int32), ("reserved", ctypes.c_uint32)]


class Interaction(ctypes.Structure):
    _fields_ = [("type", ctypes.c_int32), ("flags", ctypes.c_uint32), ("system_index", ctypes.c_int64),
                ("payload_offset", ctypes.c_uint64), ("payload_size", ctypes.c_uint64)]


assert ctypes.sizeof(Interaction) == 32


class Batch(ctypes.Structure):
    _fields_ = [("struct_size", ctypes.c_uint32), ("api_version", ctypes.c_uint32),
                ("batch_size", ctypes.c_int64), ("total_atoms", ctypes.c_int64),
                ("total_point_charges", ctypes.c_int64),
                ("total_charge_response_elements", ctypes.c_int64),
                ("atom_offsets", ConstBuffer), ("atomic_numbers", ConstBuffer),
                ("positions", ConstBuffer), ("molecular_charges", ConstBuffer),
                ("unpaired_electrons", ConstBuffer), ("point_charge_offsets", ConstBuffer),
                ("point_charge_positions", ConstBuffer), ("point_charge_values", ConstBuffer),
                ("point_charge_gammas", ConstBuffer), ("atomic_potential_shifts", ConstBuffer),
                ("charge_response_offsets", ConstBuffer), ("charge_response_matrix", ConstBuffer),
                ("spin_channels", ConstBuffer), ("total_interactions", ctypes.c_int64),
                ("interaction_descriptors", ConstBuffer), ("interaction_payload", ConstBuffer),
                ("cell_matrices", ConstBuffer), ("periodic_axes", ConstBuffer)]


class ComputeOptions(ctypes.Structure):
    _fields_ = [("struct_size", ctypes.c_uint32), ("api_version", ctypes.c_uint32),
                ("model", ctypes.c_int32), ("flags", ctypes.c_uint32),
                ("max_scc_iterations", ctypes.c_int32), ("reserved", ctypes.c_uint32),
                ("charge_tolerance", ctypes.c_double), ("energy_tolerance", ctypes.c_double),
                ("electronic_temperature", ctypes.c_double), ("scc_start_mode", ctypes.c_int32),
                ("reserved_v2", ctypes.c_uint32), ("scc_mixer", ctypes.c_int32),
                ("scc_mixer_history", ctypes.c_int32), ("scc_mixer_damping", ctypes.c_double),
                ("determinism", ctypes.c_int32), ("reserved_v3", ctypes.c_uint32)]


class BatchResult(ctypes.Structure):
    _fields_ = [("struct_size", ctypes.c_uint32), ("api_version", ctypes.c_uint32),
                ("flags", ctypes.c_uint32), ("reserved", ctypes.c_uint32),
                ("energies", Buffer), ("forces", Buffer), ("atomic_charges", Buffer),
                ("point_charge_forces", Buffer), ("scc_iterations", Buffer),
                ("scc_converged", Buffer), ("per_system_status", Buffer),
                ("dipole_moments", Buffer), ("quadrupole_moments", Buffer),
                ("wiberg_orders", Buffer), ("spin_populations", Buffer)]


KERNEL_FAMILIES = ("geometry", "integrals", "cholesky", "potentials", "fock", "reduce", "tridiag",
                   "fermi", "back", "density", "mulliken", "energy_mix", "energy_weighted",
                   "force_pairs", "force_atoms", "force_atm", "tridiag_eigen", "backtransform")


class KernelStat(ctypes.Structure):
    _fields_ = [("launches", ctypes.c_int32), ("reserved", ctypes.c_int32),
                ("total_ms", ctypes.c_double), ("flops", ctypes.c_double),
                ("bytes", ctypes.c_double)]


class Extension(ctypes.Structure):
    _fields_ = [("struct_size", ctypes.c_uint32), ("api_version", ctypes.c_uint32),
                ("point_charge_potential_gradient", ctypes.c_void_p),
                ("overlap", ctypes.c_void_p), ("core_hamiltonian", ctypes.c_void_p),
                ("density", ctypes.c_void_p), ("coordination_numbers", ctypes.c_void_p),
                ("energy_components", ctypes.c_void_p), ("kernel_stats", ctypes.c_void_p),
                ("setup_ms", ctypes.c_float),
                ("scc_ms", ctypes.c_float), ("force_ms", ctypes.c_float),
                ("total_ms", ctypes.c_float), ("scc_rounds", ctypes.c_int32),
                ("kernel_stats_mask", ctypes.c_uint32),
                ("orbital_energies", ctypes.c_void_p), ("occupations", ctypes.c_void_p)]


class RequestInfo(ctypes.Structure):
    _fields_ = [("struct_size", ctypes.c_uint32), ("api_version", ctypes.c_uint32),
                ("state", ctypes.c_int32), ("completion_status", ctypes.c_int32),
                ("result_flags", ctypes.c_uint32), ("reserved", ctypes.c_uint32)]


REQUEST_IDLE, REQUEST_PENDING, REQUEST_COMPLETE = 0, 1, 2


class ResultOwnerOptions(ctypes.Structure):  # xtbloom_result_owner_options_t
    _fields_ = [("struct_size", ctypes.c_uint32), ("api_version", ctypes.c_uint32),
                ("memory_space", ctypes.c_int32), ("device_id", ctypes.c_int32),
                ("size_bytes", ctypes.c_uint64), ("reserved", ctypes.c_uint32)]


class DLPackView(ctypes.Structure):  # xtbloom_dlpack_view_t
    _fields_ = [("struct_size", ctypes.c_uint32), ("api_version", ctypes.c_uint32),
                ("byte_offset", ctypes.c_uint64), ("dtype_code", ctypes.c_int32),
                ("dtype_bits", ctypes.c_int32), ("dtype_lanes", ctypes.c_int32),
                ("ndim", ctypes.c_int32), ("reserved", ctypes.c_uint32),
                ("shape", ctypes.POINTER(ctypes.c_int64))]


class WorkspaceQuery(ctypes.Structure):
    _fields_ = [("struct_size", ctypes.c_uint32), ("api_version", ctypes.c_uint32),
                ("compute_flags", ctypes.c_uint32), ("reserved", ctypes.c_uint32),
                ("host_required_bytes", ctypes.c_uint64),
                ("host_required_alignment", ctypes.c_uint32),
                ("device_required_bytes", ctypes.c_uint64),
                ("device_required_alignment", ctypes.c_uint32), ("reserved_v2", ctypes.c_uint32)]


# Byte images asserted by the reference ABI for LP64 (reference xtbloom.h:320-343).
assert ctypes.sizeof(ContextOptions) == 32
assert ctypes.sizeof(Buffer) == 24
assert ctypes.sizeof(Batch) == 456
assert ctypes.sizeof(ComputeOptions) == 80
assert ctypes.sizeof(BatchResult) == 280
assert ctypes.sizeof(KernelStat) == 32
assert ctypes.sizeof(WorkspaceQuery) == 48 and WorkspaceQuery.device_required_bytes.offset == 32
assert ctypes.sizeof(RequestInfo) == 24
assert ctypes.sizeof(ResultOwnerOptions) == 32 and ctypes.sizeof(DLPackView) == 48

EXPORTED_SYMBOLS = (
    "xtbloom_version_string", "xtbloom_status_string", "xtbloom_get_last_error",
    "xtbloom_context_options_init", "xtbloom_batch_init", "xtbloom_compute_options_init",
    "xtbloom_batch_result_init", "xtbloom_workspace_query_init", "xtbloom_request_info_init",
    "xtbloom_context_create", "xtbloom_context_destroy", "xtbloom_context_get_backend",
    "xtbloom_context_get_device_id", "xtbloom_compute", "xtbloom_request_create",
    "xtbloom_request_query", "xtbloom_request_wait", "xtbloom_request_get_error",
    "xtbloom_request_destroy", "xtbloom_compute_enqueue", "xtbloom_plan_create",
    "xtbloom_plan_destroy", "xtbloom_plan_query_workspace", "xtbloom_plan_compute",
    "xtbloom_plan_compute_enqueue", "xtbloom_result_owner_options_init",
    "xtbloom_result_owner_create", "xtbloom_result_owner_buffer", "xtbloom_result_owner_retain",
    "xtbloom_result_owner_release", "xtbloom_result_owner_export_dltensor",
    "xtbloom_amd_compute_extended", "xtbloom_amd_packed_matrix_elements",
    "xtbloom_amd_compute_no_alloc",
)

_LIB = None


def library_path() -> str:
    return os.environ.get("XTBLOOM_AMD_LIBRARY", DEFAULT_LIBRARY)


def load() -> ctypes.CDLL:
    """Load the HIP library once; raise if it has not been built."""
    global _LIB
    if _LIB is not None:
        return _LIB
    path = library_path()
    if not os.path.exists(path):
        raise OSError(f"{path} is missing: build it with `python -m xtbloom_amd.build` "
                      "(the ROCm backend has no CPU fallback)")
    lib = ctypes.CDLL(path)
    vp = ctypes.c_void_p
    lib.xtbloom_version_string.restype = ctypes.c_char_p
    lib.xtbloom_status_string.restype = ctypes.c_char_p
    lib.xtbloom_status_string.argtypes = [ctypes.c_int32]
    lib.xtbloom_get_last_error.restype = ctypes.c_char_p
    for name, struct in (("xtbloom_context_options_init", ContextOptions),
                         ("xtbloom_batch_init", Batch),
                         ("xtbloom_compute_options_init", ComputeOptions),
                         ("xtbloom_batch_result_init", BatchResult)):
        fn = getattr(lib, name)
        fn.argtypes = [ctypes.POINTER(struct), ctypes.c_size_t]
        fn.restype = ctypes.c_int32
    lib.xtbloom_context_create.argtypes = [ctypes.POINTER(ContextOptions), ctypes.POINTER(vp)]
    lib.xtbloom_context_create.restype = ctypes.c_int32
    lib.xtbloom_context_destroy.argtypes = [vp]
    lib.xtbloom_context_destroy.restype = None
    lib.xtbloom_context_get_backend.argtypes = [vp]
    lib.xtbloom_context_get_backend.restype = ctypes.c_int32
    lib.xtbloom_context_get_device_id.argtypes = [vp]
    lib.xtbloom_context_get_device_id.restype = ctypes.c_int32
    lib.xtbloom_compute.argtypes = [vp, ctypes.POINTER(Batch), ctypes.POINTER(ComputeOptions),
                                    ctypes.POINTER(BatchResult)]
    lib.xtbloom_compute.restype = ctypes.c_int32
    lib.xtbloom_amd_compute_extended.argtypes = [vp, ctypes.POINTER(Batch),
                                                 ctypes.POINTER(ComputeOptions),
                                                 ctypes.POINTER(BatchResult),
                                                 ctypes.POINTER(Extension)]
    lib.xtbloom_amd_compute_extended.restype = ctypes.c_int32
    lib.xtbloom_amd_compute_no_alloc.argtypes = [vp, ctypes.POINTER(Batch), ctypes.POINTER(ComputeOptions),
                                                 ctypes.POINTER(BatchResult)]
    lib.xtbloom_amd_compute_no_alloc.restype = ctypes.c_int32
    lib.xtbloom_amd_packed_matrix_elements.argtypes = [ctypes.POINTER(Batch)]
    lib.xtbloom_amd_packed_matrix_elements.restype = ctypes.c_int64
    lib.xtbloom_workspace_query_init.argtypes = [ctypes.POINTER(WorkspaceQuery), ctypes.c_size_t]
    lib.xtbloom_workspace_query_init.restype = ctypes.c_int32
    lib.xtbloom_plan_create.argtypes = [vp, ctypes.POINTER(Batch), ctypes.POINTER(ComputeOptions),
                                        ctypes.POINTER(vp)]
    lib.xtbloom_plan_create.restype = ctypes.c_int32
    lib.xtbloom_plan_destroy.argtypes = [vp]
    lib.xtbloom_plan_destroy.restype = None
    lib.xtbloom_plan_query_workspace.argtypes = [vp, ctypes.POINTER(WorkspaceQuery)]
    lib.xtbloom_plan_query_workspace.restype = ctypes.c_int32
    lib.xtbloom_plan_compute.argtypes = [vp, ctypes.POINTER(Batch), ctypes.POINTER(ComputeOptions),
                                         ctypes.POINTER(BatchResult)]
    lib.xtbloom_plan_compute.restype = ctypes.c_int32
    lib.xtbloom_request_info_init.argtypes = [ctypes.POINTER(RequestInfo), ctypes.c_size_t]
    lib.xtbloom_request_info_init.restype = ctypes.c_int32
    lib.xtbloom_request_create.argtypes = [vp, ctypes.POINTER(vp)]
    lib.xtbloom_request_create.restype = ctypes.c_int32
    for name in ("xtbloom_request_query", "xtbloom_request_wait"):
        getattr(lib, name).argtypes = [vp, ctypes.POINTER(RequestInfo)]
        getattr(lib, name).restype = ctypes.c_int32
    lib.xtbloom_request_get_error.argtypes = [vp]
    lib.xtbloom_request_get_error.restype = ctypes.c_char_p
    lib.xtbloom_request_destroy.argtypes = [vp]
    lib.xtbloom_request_destroy.restype = None
    lib.xtbloom_compute_enqueue.argtypes = [vp, ctypes.POINTER(Batch), ctypes.POINTER(ComputeOptions),
                                            ctypes.POINTER(BatchResult), vp]
    lib.xtbloom_compute_enqueue.restype = ctypes.c_int32
    lib.xtbloom_plan_compute_enqueue.argtypes = [vp, ctypes.POINTER(Batch),
                                                 ctypes.POINTER(ComputeOptions),
                                                 ctypes.POINTER(BatchResult), vp]
    lib.xtbloom_plan_compute_enqueue.restype = ctypes.c_int32
    lib.xtbloom_result_owner_options_init.argtypes = [ctypes.POINTER(ResultOwnerOptions), ctypes.c_size_t]
    lib.xtbloom_result_owner_options_init.restype = ctypes.c_int32
    lib.xtbloom_result_owner_create.argtypes = [ctypes.POINTER(ResultOwnerOptions), ctypes.POINTER(vp)]
    lib.xtbloom_result_owner_create.restype = ctypes.c_int32
    lib.xtbloom_result_owner_buffer.argtypes = [vp, ctypes.POINTER(Buffer)]
    lib.xtbloom_result_owner_buffer.restype = ctypes.c_int32
    lib.xtbloom_result_owner_retain.argtypes = [vp]
    lib.xtbloom_result_owner_retain.restype = None
    lib.xtbloom_result_owner_release.argtypes = [vp]
    lib.xtbloom_result_owner_release.restype = None
    lib.xtbloom_result_owner_export_dltensor.argtypes = [vp, ctypes.POINTER(DLPackView), ctypes.c_int,
                                                         ctypes.POINTER(vp)]
    lib.xtbloom_result_owner_export_dltensor.restype = ctypes.c_int32
    _LIB = lib
    return lib


def last_error() -> str:
    return load().xtbloom_get_last_error().decode()


def status_string(status: int) -> str:
    return load().xtbloom_status_string(status).decode()
