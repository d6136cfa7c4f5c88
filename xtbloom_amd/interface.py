"""Python host layer of the MI355X backend, mirroring xTBloom's ``Calculator`` API.

The classes keep the reference's names, argument meaning and result keys
(reference python/xtbloom/interface.py: PointCharge :187, Structure :291,
Context :500, Result :1433, BatchResult :1506, Calculator :1792,
BatchCalculator :2028) with the backend map extended by ``"hip"`` ->
``BACKEND_ROCM`` (reference interface.py:154-158 has auto/cpu/cuda).  Every
call goes through ``xtbloom_compute`` of ``libxtbloom_amd.so``; there is no
Python or CPU compute path.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass
from typing import Sequence

import numpy as np

from . import library

_METHODS = {"GFN2-XTB": library.MODEL_GFN2_XTB, "GFN2": library.MODEL_GFN2_XTB,
            "GFN1-XTB": library.MODEL_GFN1_XTB, "GFN1": library.MODEL_GFN1_XTB}
_BACKENDS = {"auto": library.BACKEND_AUTO, "cpu": library.BACKEND_CPU,
             "cuda": library.BACKEND_CUDA, "hip": library.BACKEND_ROCM,
             "rocm": library.BACKEND_ROCM}


class XTBloomError(RuntimeError):
    """A non-success status from the C ABI."""

    def __init__(self, status: int, message: str):
        super().__init__(f"{library.status_string(status)}: {message}")
        self.status = status


class XTBloomValueError(ValueError):
    pass


def _check(status: int):
    if status != library.STATUS_SUCCESS:
        raise XTBloomError(status, library.last_error())


@dataclass
class PointCharge:
    """External point charges (positions bohr, charges e, gammas Hartree)."""
    positions: np.ndarray
    charges: np.ndarray
    gammas: np.ndarray

    def __post_init__(self):
        self.positions = np.ascontiguousarray(self.positions, dtype=float)
        self.charges = np.ascontiguousarray(self.charges, dtype=float)
        self.gammas = np.ascontiguousarray(self.gammas, dtype=float)
        if self.positions.ndim != 2 or self.positions.shape[1] != 3:
            raise XTBloomValueError("point charge positions must have shape (n, 3)")
        n = len(self.positions)
        if self.charges.shape != (n,) or self.gammas.shape != (n,) or n == 0:
            raise XTBloomValueError("point charge values/gammas must match a nonempty position set")
        if not (np.isfinite(self.positions).all() and np.isfinite(self.charges).all()
                and np.isfinite(self.gammas).all()):
            raise XTBloomValueError("point charge inputs must be finite")


@dataclass(frozen=True)
class ChargeResponse:
    """External charge operator of one molecule (periodic QM/MM coupling,
    reference python/xtbloom/interface.py:231-267): an SCC shift ``b + A q``
    on the atomic charges and the energy ``q.b + q.A.q / 2``.  ``shifts`` is b
    (n, Hartree/e), ``matrix`` the symmetric A (n x n, Hartree/e^2); both are
    constant operators, so forces leave out their coordinate derivatives."""
    shifts: np.ndarray
    matrix: np.ndarray

    def __post_init__(self):
        shifts = np.asarray(self.shifts, dtype=float)
        matrix = np.asarray(self.matrix, dtype=float)
        if shifts.ndim != 1:
            raise XTBloomValueError("charge response shifts must be one-dimensional")
        if matrix.ndim != 2 or matrix.shape[0] != matrix.shape[1]:
            raise XTBloomValueError("charge response matrix must be square")
        if matrix.shape[0] != shifts.size:
            raise XTBloomValueError("charge response matrix must match the shift count")
        if not (np.isfinite(shifts).all() and np.isfinite(matrix).all()):
            raise XTBloomValueError("charge response inputs must be finite")
        object.__setattr__(self, "shifts", np.ascontiguousarray(shifts))
        object.__setattr__(self, "matrix", np.ascontiguousarray(matrix))


def _normalize_efield(efield):
    """A uniform electric field (Hartree per elementary charge per bohr) as a
    frozen length-3 vector; ``None`` and the zero vector mean no attachment
    (reference python/xtbloom/interface.py:273-289)."""
    if efield is None:
        return None
    value = np.asarray(efield, dtype=float)
    if value.ndim != 1 or value.shape[0] != 3:
        raise XTBloomValueError("efield must be a vector of length 3")
    if not np.isfinite(value).all():
        raise XTBloomValueError("efield must be finite")
    if np.any(value != 0.0):
        value = np.array(value)
        value.flags.writeable = False
        return value
    return None


class Structure:
    """One molecule: atomic numbers, positions (bohr), charge, uhf, optional
    point charges and an optional uniform electric field (atomic units).
    ``spin_channels`` 1 (restricted) or 2 (unrestricted) defaults to 2 for an
    open shell (uhf != 0), as in the reference (python/xtbloom/interface.py:
    158-162, 350-371)."""

    def __init__(self, numbers, positions, charge: float = 0.0, uhf: int | None = None,
                 point_charges: PointCharge | None = None, efield=None,
                 charge_response: ChargeResponse | None = None, spin_channels: int | None = None):
        self._numbers = np.ascontiguousarray(numbers, dtype=np.int32)
        self._positions = np.ascontiguousarray(positions, dtype=float).reshape(-1, 3)
        if self._numbers.ndim != 1 or len(self._numbers) != len(self._positions):
            raise XTBloomValueError("numbers and positions must describe the same atoms")
        self._charge = float(charge)
        self._uhf = int(uhf or 0)
        spin = (2 if self._uhf != 0 else 1) if spin_channels is None else int(spin_channels)
        if spin not in (1, 2):
            raise XTBloomValueError("spin_channels must be 1 (restricted) or 2 (unrestricted)")
        self._spin_channels = spin
        self._point_charges = point_charges
        self._efield = _normalize_efield(efield)
        if charge_response is not None and charge_response.shifts.size != len(self._numbers):
            raise XTBloomValueError("charge response shifts must match the atom count")
        self._charge_response = charge_response

    numbers = property(lambda self: self._numbers)
    positions = property(lambda self: self._positions)
    charge = property(lambda self: self._charge)
    uhf = property(lambda self: self._uhf)
    point_charges = property(lambda self: self._point_charges)
    efield = property(lambda self: self._efield)
    spin_channels = property(lambda self: self._spin_channels)
    charge_response = property(lambda self: self._charge_response)

    def update(self, positions=None, point_charges=None):
        if positions is not None:
            pos = np.ascontiguousarray(positions, dtype=float).reshape(-1, 3)
            if pos.shape != self._positions.shape:
                raise XTBloomValueError("updated positions must keep the atom count")
            self._positions = pos
        if point_charges is not None:
            self._point_charges = point_charges


class Context:
    """Owns one native context (one device, one stream)."""

    def __init__(self, backend: str | int = "hip", device_id: int | None = None, stream=None):
        lib = library.load()
        self.backend_code = _BACKENDS[backend] if isinstance(backend, str) else int(backend)
        opts = library.ContextOptions()
        _check(lib.xtbloom_context_options_init(ctypes.byref(opts), ctypes.sizeof(opts)))
        opts.backend = self.backend_code
        opts.device_id = -1 if device_id is None else int(device_id)
        opts.stream = stream
        handle = ctypes.c_void_p()
        _check(lib.xtbloom_context_create(ctypes.byref(opts), ctypes.byref(handle)))
        self._handle = handle

    @property
    def handle(self):
        return self._handle

    @property
    def backend(self) -> int:
        return library.load().xtbloom_context_get_backend(self._handle)

    @property
    def device_id(self) -> int:
        return library.load().xtbloom_context_get_device_id(self._handle)

    def close(self):
        if getattr(self, "_handle", None):
            library.load().xtbloom_context_destroy(self._handle)
            self._handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _const(arr, space=library.MEMORY_HOST):
    b = library.ConstBuffer()
    if arr is None:
        return b
    b.data = arr.ctypes.data
    b.size_bytes = arr.nbytes
    b.memory_space = space
    return b


def _out(arr):
    b = library.Buffer()
    if arr is None:
        return b
    b.data = arr.ctypes.data
    b.size_bytes = arr.nbytes
    b.memory_space = library.MEMORY_HOST
    return b


@dataclass
class ComputedBatch:
    """Flat outputs of one native call (forces are -dE/dR)."""
    atom_offsets: np.ndarray
    point_offsets: np.ndarray | None
    energies: np.ndarray
    forces: np.ndarray
    charges: np.ndarray
    point_charge_forces: np.ndarray | None
    dipole_moments: np.ndarray | None
    scc_iterations: np.ndarray
    scc_converged: np.ndarray
    per_system_status: np.ndarray
    point_charge_potential_gradient: np.ndarray | None = None
    extension: dict | None = None
    result_flags: int = 0  # xtbloom_batch_result_t.flags of the call


_DEFAULT_FLAGS = library.COMPUTE_ENERGY | library.COMPUTE_FORCES | library.COMPUTE_ATOMIC_CHARGES


def _field_attachments(B: int, electric_field):
    """Interaction descriptors and the payload for uniform-field attachments
    (reference python/xtbloom/interface.py:1172-1184).  ``electric_field`` is a
    (B, 3) array (every system attached, zero rows included) or a sequence of B
    entries, each None (no attachment) or a length-3 vector."""
    rows = list(electric_field)
    if len(rows) != B:
        raise XTBloomValueError("electric_field needs one entry per system")
    attached = [(i, np.asarray(r, dtype=float)) for i, r in enumerate(rows) if r is not None]
    desc = (library.Interaction * max(1, len(attached)))()
    payload = np.zeros(max(1, len(attached)) * library.ELECTRIC_FIELD_BLOCK_BYTES // 8, np.float64)
    head = payload.view(np.int32)
    for k, (i, field) in enumerate(attached):
        if field.shape != (3,):
            raise XTBloomValueError("an electric field must be a vector of length 3")
        desc[k].type = library.INTERACTION_ELECTRIC_FIELD
        desc[k].flags = 0
        desc[k].system_index = i
        desc[k].payload_offset = k * library.ELECTRIC_FIELD_BLOCK_BYTES
        desc[k].payload_size = library.ELECTRIC_FIELD_BLOCK_BYTES
        head[8 * k] = 1  # block_version
        head[8 * k + 1] = 0  # reserved
        payload[4 * k + 1:4 * k + 4] = field
    return len(attached), desc, payload


def _descriptors(lib, atom_offsets, numbers, positions, charges, uhf, point_offsets,
                 point_positions, point_values, point_gammas, model, flags, max_scc_iterations,
                 charge_tolerance, energy_tolerance, electronic_temperature, scc_mixer_history,
                 scc_mixer_damping, scc_start_mode, determinism=0, electric_field=None,
                 atomic_potential_shifts=None, charge_response_offsets=None,
                 charge_response_matrix=None, spin_channels=None):
    """Batch and compute-options descriptors over flat host arrays (kept alive by the caller)."""
    atom_offsets = np.ascontiguousarray(atom_offsets, dtype=np.int64)
    numbers = np.ascontiguousarray(numbers, dtype=np.int32)
    positions = np.ascontiguousarray(positions, dtype=np.float64).reshape(-1, 3)
    B, N = len(atom_offsets) - 1, len(numbers)
    charges = np.zeros(B) if charges is None else np.ascontiguousarray(charges, dtype=np.float64)
    uhf = np.zeros(B, np.int32) if uhf is None else np.ascontiguousarray(uhf, dtype=np.int32)
    P = 0
    if point_offsets is not None:
        point_offsets = np.ascontiguousarray(point_offsets, dtype=np.int64)
        point_positions = np.ascontiguousarray(point_positions, dtype=np.float64).reshape(-1, 3)
        point_values = np.ascontiguousarray(point_values, dtype=np.float64)
        point_gammas = np.ascontiguousarray(point_gammas, dtype=np.float64)
        P = len(point_values)
        if P == 0:
            point_offsets = point_positions = point_values = point_gammas = None
    batch = library.Batch()
    _check(lib.xtbloom_batch_init(ctypes.byref(batch), ctypes.sizeof(batch)))
    batch.batch_size, batch.total_atoms, batch.total_point_charges = B, N, P
    batch.atom_offsets = _const(atom_offsets)
    batch.atomic_numbers = _const(numbers)
    batch.positions = _const(positions)
    batch.molecular_charges = _const(charges)
    batch.unpaired_electrons = _const(uhf)
    if P:
        batch.point_charge_offsets = _const(point_offsets)
        batch.point_charge_positions = _const(point_positions)
        batch.point_charge_values = _const(point_values)
        batch.point_charge_gammas = _const(point_gammas)
    fields = None
    if electric_field is not None:
        count, desc, payload = _field_attachments(B, electric_field)
        if count:
            fields = (desc, payload)
            batch.total_interactions = count
            batch.interaction_descriptors = library.ConstBuffer(ctypes.addressof(desc), ctypes.sizeof(desc),
                                                                library.MEMORY_HOST, 0)
            batch.interaction_payload = _const(payload)
    operator = []
    if spin_channels is not None:  # [B] 1 or 2 (absent: every system restricted)
        operator.append(np.ascontiguousarray(spin_channels, dtype=np.int32))
        batch.spin_channels = _const(operator[-1])
    if atomic_potential_shifts is not None:
        operator.append(np.ascontiguousarray(atomic_potential_shifts, dtype=np.float64))
        batch.atomic_potential_shifts = _const(operator[-1])
    if charge_response_matrix is not None or charge_response_offsets is not None:
        roff = np.ascontiguousarray(charge_response_offsets, dtype=np.int64)
        rmat = np.ascontiguousarray(charge_response_matrix, dtype=np.float64)
        operator += [roff, rmat]
        batch.total_charge_response_elements = rmat.size
        batch.charge_response_offsets = _const(roff)
        batch.charge_response_matrix = _const(rmat)
    opts = library.ComputeOptions()
    _check(lib.xtbloom_compute_options_init(ctypes.byref(opts), ctypes.sizeof(opts)))
    opts.model = model
    opts.flags = flags if P else flags & ~library.COMPUTE_POINT_CHARGE_FORCES
    opts.max_scc_iterations = max_scc_iterations
    opts.charge_tolerance = charge_tolerance
    opts.energy_tolerance = energy_tolerance
    opts.electronic_temperature = electronic_temperature * library.KELVIN_TO_HARTREE
    opts.scc_mixer_history = scc_mixer_history
    opts.scc_mixer_damping = scc_mixer_damping
    opts.scc_start_mode = scc_start_mode
    opts.determinism = determinism
    keep = (atom_offsets, numbers, positions, charges, uhf, point_offsets, point_positions,
            point_values, point_gammas, fields, operator)
    return batch, opts, keep, (B, N, P)


def compute_arrays(context: Context, atom_offsets, numbers, positions, charges=None, uhf=None,
                   point_offsets=None, point_positions=None, point_values=None,
                   point_gammas=None, *, model=library.MODEL_GFN2_XTB, flags=_DEFAULT_FLAGS,
                   max_scc_iterations=250, charge_tolerance=1e-6, energy_tolerance=1e-8,
                   electronic_temperature=300.0, scc_mixer_history=8, scc_mixer_damping=0.4,
                   scc_start_mode=library.SCC_START_FRESH, plan: "Plan | None" = None,
                   request: "Request | None" = None,
                   extension=False, diagnostics=False, kernel_stats=False,
                   determinism=library.DETERMINISM_DEFAULT, options_hook=None,
                   no_alloc=False, wait=True, electric_field=None, atomic_potential_shifts=None,
                   charge_response_offsets=None, charge_response_matrix=None,
                   spin_channels=None) -> "ComputedBatch | PendingBatch":
    """One ``xtbloom_compute`` over flat ragged arrays (the ``xtbloom_batch_t`` layout).

    ``scc_start_mode=library.SCC_START_WARM`` starts from the converged state of
    the latest fully converged call with the same topology and policy (strict:
    no such call raises ``XTBloomError`` with INVALID_ARGUMENT).  With ``plan``
    the call goes through ``xtbloom_plan_compute`` of that plan.
    ``determinism=library.DETERMINISM_REPRODUCIBLE`` asks for bit-identical
    results on repeated calls.  ``options_hook`` (optional) receives the
    compute-options struct right before the call (tests of the validation).
    ``no_alloc=True`` goes through ``xtbloom_amd_compute_no_alloc`` (no device
    allocation; the batch must fit its fixed workspace).  With ``request`` the
    call is submitted through ``xtbloom_compute_enqueue`` (or the plan's
    enqueue); ``wait=False`` returns a :class:`PendingBatch` at once, whose
    ``result()`` waits for the request.  ``electric_field`` attaches uniform
    fields (a (B, 3) array, or B entries of None or a length-3 vector) as
    ABI-v3 interaction attachments.  ``atomic_potential_shifts`` (b, [N]),
    ``charge_response_offsets`` ([B + 1]) and ``charge_response_matrix`` (A,
    n_atoms^2 per system, packed) add the external charge operator.
    ``spin_channels`` ([B], 1 or 2) selects unrestricted systems.
    """
    lib = library.load()
    batch, opts, keep, (B, N, P) = _descriptors(
        lib, atom_offsets, numbers, positions, charges, uhf, point_offsets, point_positions,
        point_values, point_gammas, model, flags, max_scc_iterations, charge_tolerance,
        energy_tolerance, electronic_temperature, scc_mixer_history, scc_mixer_damping,
        scc_start_mode, determinism, electric_field, atomic_potential_shifts,
        charge_response_offsets, charge_response_matrix, spin_channels)
    if options_hook is not None:
        options_hook(opts)
    atom_offsets, point_offsets = keep[0], keep[5]
    energies = np.full(B, np.nan)
    forces = np.full((N, 3), np.nan)
    q = np.full(N, np.nan)
    pcf = np.full((P, 3), np.nan) if P else None
    dip = np.full((B, 3), np.nan) if flags & library.COMPUTE_DIPOLE_MOMENTS else None
    iters = np.zeros(B, np.int32)
    conv = np.zeros(B, np.uint8)
    status = np.zeros(B, np.int32)
    res = library.BatchResult()
    _check(lib.xtbloom_batch_result_init(ctypes.byref(res), ctypes.sizeof(res)))
    res.energies = _out(energies)
    res.forces = _out(forces)
    res.atomic_charges = _out(q)
    res.point_charge_forces = _out(pcf)
    res.dipole_moments = _out(dip)
    res.scc_iterations = _out(iters)
    res.scc_converged = _out(conv)
    res.per_system_status = _out(status)
    ext_out = None
    dedq = None
    if no_alloc and (request is not None or plan is not None or extension or diagnostics
                     or kernel_stats):
        raise XTBloomValueError("no-allocation calls take no plan, request or extension outputs")
    if no_alloc:
        _check(lib.xtbloom_amd_compute_no_alloc(context.handle, ctypes.byref(batch),
                                                ctypes.byref(opts), ctypes.byref(res)))
    elif request is not None:
        # xtbloom_compute_enqueue / xtbloom_plan_compute_enqueue; the request
        # reports the completion status (admission errors raise here)
        if extension or diagnostics or kernel_stats:
            raise XTBloomValueError("enqueued calls take no extension outputs")
        if plan is not None:
            _check(lib.xtbloom_plan_compute_enqueue(plan.handle, ctypes.byref(batch),
                                                    ctypes.byref(opts), ctypes.byref(res),
                                                    request.handle))
        else:
            _check(lib.xtbloom_compute_enqueue(context.handle, ctypes.byref(batch),
                                               ctypes.byref(opts), ctypes.byref(res),
                                               request.handle))
        computed = ComputedBatch(atom_offsets, point_offsets, energies, forces, q, pcf, dip, iters,
                                 conv, status, None, None)
        pending = PendingBatch(request, computed)
        return pending if not wait else pending.result()
    elif plan is not None:
        if extension or diagnostics or kernel_stats:
            raise XTBloomValueError("plan calls take no extension outputs")
        _check(lib.xtbloom_plan_compute(plan.handle, ctypes.byref(batch), ctypes.byref(opts),
                                        ctypes.byref(res)))
    elif extension or diagnostics or kernel_stats:
        ext = library.Extension()
        ext.struct_size = ctypes.sizeof(ext)
        ext.api_version = library.API_VERSION
        bufs = {}
        if P:
            dedq = np.full(P, np.nan)
            ext.point_charge_potential_gradient = dedq.ctypes.data
        if diagnostics:
            nn = int(lib.xtbloom_amd_packed_matrix_elements(ctypes.byref(batch)))
            for name in ("overlap", "core_hamiltonian", "density"):
                bufs[name] = np.zeros(nn)
                setattr(ext, name, bufs[name].ctypes.data)
            bufs["coordination_numbers"] = np.zeros(N)
            ext.coordination_numbers = bufs["coordination_numbers"].ctypes.data
            bufs["energy_components"] = np.zeros((B, 8))
            ext.energy_components = bufs["energy_components"].ctypes.data
            # packed per system over its n levels; 9 AOs per atom bound the total
            for name in ("orbital_energies", "occupations"):
                bufs[name] = np.full(9 * N, np.nan)
                setattr(ext, name, bufs[name].ctypes.data)
        stats = (library.KernelStat * len(library.KERNEL_FAMILIES))() if kernel_stats else None
        if stats is not None:
            ext.kernel_stats = ctypes.addressof(stats)
            ext.kernel_stats_mask = _stats_mask(kernel_stats)
        _check(lib.xtbloom_amd_compute_extended(context.handle, ctypes.byref(batch),
                                                ctypes.byref(opts), ctypes.byref(res),
                                                ctypes.byref(ext)))
        ext_out = dict(bufs, setup_ms=ext.setup_ms, scc_ms=ext.scc_ms, force_ms=ext.force_ms,
                       total_ms=ext.total_ms, scc_rounds=ext.scc_rounds)
        if stats is not None:
            ext_out["kernel_stats"] = _stats_dict(stats)
    else:
        _check(lib.xtbloom_compute(context.handle, ctypes.byref(batch), ctypes.byref(opts),
                                   ctypes.byref(res)))
    return ComputedBatch(atom_offsets, point_offsets, energies, forces, q, pcf, dip, iters,
                         conv.astype(bool), status, dedq, ext_out, int(res.flags))


class PendingBatch:
    """A submitted call: its request and the output arrays the library fills.

    ``done()`` polls the request without blocking; ``result()`` waits for it,
    raises ``XTBloomError`` with the completion status of a failed call, and
    returns the :class:`ComputedBatch`.
    """

    def __init__(self, request: "Request", computed: ComputedBatch):
        self.request = request
        self._computed = computed
        self._result = None

    def done(self) -> bool:
        return self.request.query()["state"] == library.REQUEST_COMPLETE

    def result(self) -> ComputedBatch:
        if self._result is None:
            info = self.request.wait()
            if info["completion_status"] != library.STATUS_SUCCESS:
                raise XTBloomError(info["completion_status"], self.request.error())
            c = self._computed
            c.scc_converged = c.scc_converged.astype(bool)
            c.result_flags = int(info.get("result_flags", 0))
            self._result = c
        return self._result


class Request:
    """Reusable submission owner (``xtbloom_request_create``) bound to a context.

    An accepted submission is PENDING until the context's worker has run it
    and published its outputs, then COMPLETE (reference ``xtbloom.h`` request
    API); submissions on one context run in order, and a synchronous call
    waits for the ones queued before it.  Close a request before its context.
    """

    def __init__(self, context: Context):
        lib = library.load()
        handle = ctypes.c_void_p()
        _check(lib.xtbloom_request_create(context.handle, ctypes.byref(handle)))
        self._handle = handle
        self._context = context

    @property
    def handle(self):
        if not self._handle:
            raise XTBloomError(library.STATUS_INVALID_ARGUMENT, "request is closed")
        return self._handle

    def _info(self, fn) -> dict:
        lib = library.load()
        info = library.RequestInfo()
        _check(lib.xtbloom_request_info_init(ctypes.byref(info), ctypes.sizeof(info)))
        _check(fn(self.handle, ctypes.byref(info)))
        return dict(state=info.state, completion_status=info.completion_status,
                    result_flags=info.result_flags)

    def query(self) -> dict:
        return self._info(library.load().xtbloom_request_query)

    def wait(self) -> dict:
        return self._info(library.load().xtbloom_request_wait)

    def error(self) -> str:
        msg = library.load().xtbloom_request_get_error(self.handle)
        return msg.decode() if msg else ""

    def close(self):
        if getattr(self, "_handle", None):
            library.load().xtbloom_request_destroy(self._handle)
            self._handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Plan:
    """Fixed-topology plan (``xtbloom_plan_create``) over flat ragged arrays.

    The plan binds the topology (atom offsets, element numbers, molecular
    charges, unpaired electrons, point-charge offsets and gammas) and the
    compute policy, and reserves the device workspace; :meth:`compute` takes
    new geometry only (reference ``python/xtbloom/library.py`` ``Plan``).  A
    plan keeps its own WARM checkpoint.  Close it before its context.
    """

    def __init__(self, context: Context, atom_offsets, numbers, positions, charges=None,
                 uhf=None, point_offsets=None, point_positions=None, point_values=None,
                 point_gammas=None, *, model=library.MODEL_GFN2_XTB, flags=_DEFAULT_FLAGS,
                 max_scc_iterations=250, charge_tolerance=1e-6, energy_tolerance=1e-8,
                 electronic_temperature=300.0, scc_mixer_history=8, scc_mixer_damping=0.4,
                 atomic_potential_shifts=None, charge_response_offsets=None,
                 charge_response_matrix=None, spin_channels=None):
        lib = library.load()
        # the operator's presence belongs to the plan; its values are inputs
        # of each call (defaulting to these)
        self._operator = dict(atomic_potential_shifts=atomic_potential_shifts,
                              charge_response_offsets=charge_response_offsets,
                              charge_response_matrix=charge_response_matrix)
        self._policy = dict(model=model, flags=flags, max_scc_iterations=max_scc_iterations,
                            charge_tolerance=charge_tolerance, energy_tolerance=energy_tolerance,
                            electronic_temperature=electronic_temperature,
                            scc_mixer_history=scc_mixer_history,
                            scc_mixer_damping=scc_mixer_damping)
        batch, opts, keep, _ = _descriptors(
            lib, atom_offsets, numbers, positions, charges, uhf, point_offsets, point_positions,
            point_values, point_gammas, scc_start_mode=library.SCC_START_FRESH, **self._policy,
            **self._operator, spin_channels=spin_channels)
        self._topology = dict(atom_offsets=keep[0], numbers=keep[1], charges=keep[3],
                              uhf=keep[4], point_offsets=keep[5], point_gammas=keep[8],
                              spin_channels=spin_channels)
        self._flags = int(opts.flags)
        handle = ctypes.c_void_p()
        _check(lib.xtbloom_plan_create(context.handle, ctypes.byref(batch), ctypes.byref(opts),
                                       ctypes.byref(handle)))
        self._handle = handle
        self._context = context  # the plan must not outlive its context

    @property
    def handle(self):
        if not self._handle:
            raise XTBloomError(library.STATUS_INVALID_ARGUMENT, "fixed-topology plan is closed")
        return self._handle

    def compute(self, positions, point_positions=None, point_values=None, *,
                scc_start_mode=library.SCC_START_FRESH, electric_field=None,
                atomic_potential_shifts=None, charge_response_matrix=None) -> ComputedBatch:
        """One ``xtbloom_plan_compute`` with new geometry (and, optionally, new
        uniform fields: a plan accepts any field, a WARM start only the
        checkpoint's; and new b or A values for a plan made with them)."""
        op = dict(self._operator)
        if atomic_potential_shifts is not None:
            op["atomic_potential_shifts"] = atomic_potential_shifts
        if charge_response_matrix is not None:
            op["charge_response_matrix"] = charge_response_matrix
        return compute_arrays(self._context, positions=positions, point_positions=point_positions,
                              point_values=point_values, scc_start_mode=scc_start_mode, plan=self,
                              electric_field=electric_field, **self._topology, **self._policy, **op)

    def query_workspace(self) -> dict:
        """Reserved workspace (``xtbloom_plan_query_workspace``)."""
        lib = library.load()
        q = library.WorkspaceQuery()
        _check(lib.xtbloom_workspace_query_init(ctypes.byref(q), ctypes.sizeof(q)))
        q.compute_flags = self._flags
        _check(lib.xtbloom_plan_query_workspace(self.handle, ctypes.byref(q)))
        return dict(host_bytes=q.host_required_bytes, host_alignment=q.host_required_alignment,
                    device_bytes=q.device_required_bytes,
                    device_alignment=q.device_required_alignment)

    def close(self):
        if getattr(self, "_handle", None):
            library.load().xtbloom_plan_destroy(self._handle)
            self._handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _stats_mask(kernel_stats) -> int:
    """``True``: events for every kernel family; an iterable of family names: those only."""
    if kernel_stats is True:
        return 0
    mask = 0
    for name in kernel_stats:
        mask |= 1 << library.KERNEL_FAMILIES.index(name)
    return mask


def _stats_dict(stats) -> dict:
    return {name: dict(launches=st.launches, ms=st.total_ms, flops=st.flops, bytes=st.bytes)
            for name, st in zip(library.KERNEL_FAMILIES, stats)}


def _tensor_buffer(cls, tensor):
    b = cls()
    if tensor is None:
        return b
    if not tensor.is_contiguous():
        raise XTBloomValueError("tensor buffers must be contiguous")
    b.data = tensor.data_ptr()
    b.size_bytes = tensor.numel() * tensor.element_size()
    b.memory_space = library.MEMORY_ROCM_DEVICE if tensor.is_cuda else library.MEMORY_HOST
    return b


def compute_tensors(context: Context, atom_offsets, numbers, positions, charges, uhf,
                    point_offsets=None, point_positions=None, point_values=None,
                    point_gammas=None, *, flags=library.COMPUTE_ENERGY | library.COMPUTE_FORCES
                    | library.COMPUTE_ATOMIC_CHARGES, max_scc_iterations=250,
                    charge_tolerance=1e-6, energy_tolerance=1e-8, electronic_temperature=300.0,
                    scc_mixer_history=8, scc_mixer_damping=0.4, outputs=None,
                    point_charge_potential_gradient=False, kernel_stats=False,
                    scc_start_mode=library.SCC_START_FRESH, plan: "Plan | None" = None):
    """``xtbloom_compute`` on torch tensors (device-resident inputs and outputs stay on the GPU).

    Inputs follow the ``xtbloom_batch_t`` dtypes (int64 offsets, int32 numbers and
    unpaired counts, float64 coordinates).  ``outputs`` may pass preallocated
    tensors (keys energies, forces, charges, point_charge_forces, scc_iterations,
    scc_converged, per_system_status, point_charge_potential_gradient); missing
    ones are allocated on the positions' device.  Returns (outputs, kernel stats).
    """
    import torch

    lib = library.load()
    dev = positions.device
    B, N = atom_offsets.numel() - 1, numbers.numel()
    P = 0 if point_values is None else point_values.numel()
    batch = library.Batch()
    _check(lib.xtbloom_batch_init(ctypes.byref(batch), ctypes.sizeof(batch)))
    batch.batch_size, batch.total_atoms, batch.total_point_charges = B, N, P
    batch.atom_offsets = _tensor_buffer(library.ConstBuffer, atom_offsets)
    batch.atomic_numbers = _tensor_buffer(library.ConstBuffer, numbers)
    batch.positions = _tensor_buffer(library.ConstBuffer, positions)
    batch.molecular_charges = _tensor_buffer(library.ConstBuffer, charges)
    batch.unpaired_electrons = _tensor_buffer(library.ConstBuffer, uhf)
    if P:
        batch.point_charge_offsets = _tensor_buffer(library.ConstBuffer, point_offsets)
        batch.point_charge_positions = _tensor_buffer(library.ConstBuffer, point_positions)
        batch.point_charge_values = _tensor_buffer(library.ConstBuffer, point_values)
        batch.point_charge_gammas = _tensor_buffer(library.ConstBuffer, point_gammas)
    opts = library.ComputeOptions()
    _check(lib.xtbloom_compute_options_init(ctypes.byref(opts), ctypes.sizeof(opts)))
    opts.flags = flags if P else flags & ~library.COMPUTE_POINT_CHARGE_FORCES
    opts.max_scc_iterations = max_scc_iterations
    opts.charge_tolerance = charge_tolerance
    opts.energy_tolerance = energy_tolerance
    opts.electronic_temperature = electronic_temperature * library.KELVIN_TO_HARTREE
    opts.scc_mixer_history = scc_mixer_history
    opts.scc_mixer_damping = scc_mixer_damping
    opts.scc_start_mode = scc_start_mode
    out = dict(outputs or {})
    f64 = dict(dtype=torch.float64, device=dev)
    out.setdefault("energies", torch.empty(B, **f64))
    if flags & library.COMPUTE_FORCES:
        out.setdefault("forces", torch.empty((N, 3), **f64))
    if flags & library.COMPUTE_ATOMIC_CHARGES:
        out.setdefault("charges", torch.empty(N, **f64))
    if P and flags & library.COMPUTE_POINT_CHARGE_FORCES:
        out.setdefault("point_charge_forces", torch.empty((P, 3), **f64))
    out.setdefault("scc_iterations", torch.empty(B, dtype=torch.int32, device=dev))
    out.setdefault("scc_converged", torch.empty(B, dtype=torch.uint8, device=dev))
    out.setdefault("per_system_status", torch.empty(B, dtype=torch.int32, device=dev))
    res = library.BatchResult()
    _check(lib.xtbloom_batch_result_init(ctypes.byref(res), ctypes.sizeof(res)))
    res.energies = _tensor_buffer(library.Buffer, out.get("energies"))
    res.forces = _tensor_buffer(library.Buffer, out.get("forces"))
    res.atomic_charges = _tensor_buffer(library.Buffer, out.get("charges"))
    res.point_charge_forces = _tensor_buffer(library.Buffer, out.get("point_charge_forces"))
    res.scc_iterations = _tensor_buffer(library.Buffer, out["scc_iterations"])
    res.scc_converged = _tensor_buffer(library.Buffer, out["scc_converged"])
    res.per_system_status = _tensor_buffer(library.Buffer, out["per_system_status"])
    ext = library.Extension()
    ext.struct_size = ctypes.sizeof(ext)
    ext.api_version = library.API_VERSION
    host_dedq = None
    if P and point_charge_potential_gradient:
        # the extension's outputs are host buffers
        host_dedq = np.empty(P)
        ext.point_charge_potential_gradient = host_dedq.ctypes.data
    stats = (library.KernelStat * len(library.KERNEL_FAMILIES))() if kernel_stats else None
    if stats is not None:
        ext.kernel_stats = ctypes.addressof(stats)
        ext.kernel_stats_mask = _stats_mask(kernel_stats)
    if dev.type == "cuda":
        torch.cuda.synchronize(dev)
    if plan is not None:
        # plan calls carry no extension: no dE/dQ and no kernel statistics
        if host_dedq is not None or stats is not None:
            raise XTBloomValueError("plan calls take no extension outputs")
        _check(lib.xtbloom_plan_compute(plan.handle, ctypes.byref(batch), ctypes.byref(opts),
                                        ctypes.byref(res)))
        return out, dict(setup_ms=0.0, scc_ms=0.0, force_ms=0.0, total_ms=0.0, scc_rounds=0)
    _check(lib.xtbloom_amd_compute_extended(context.handle, ctypes.byref(batch),
                                            ctypes.byref(opts), ctypes.byref(res),
                                            ctypes.byref(ext)))
    if host_dedq is not None:
        out["point_charge_potential_gradient"] = torch.from_numpy(host_dedq).to(dev)
    timing = dict(setup_ms=ext.setup_ms, scc_ms=ext.scc_ms, force_ms=ext.force_ms,
                  total_ms=ext.total_ms, scc_rounds=ext.scc_rounds)
    if stats is not None:
        timing["kernel_stats"] = _stats_dict(stats)
    return out, timing


class Result:
    """Single-system results; keys as in the reference (``gradient`` = -forces)."""

    def __init__(self, computed: ComputedBatch, index: int = 0):
        b, e = int(computed.atom_offsets[index]), int(computed.atom_offsets[index + 1])
        self.energy = float(computed.energies[index])
        self.forces = computed.forces[b:e].copy()
        self.charges = computed.charges[b:e].copy()
        self.point_charge_forces = None
        self.point_charge_potential_gradient = None
        if computed.point_charge_forces is not None:
            pb, pe = int(computed.point_offsets[index]), int(computed.point_offsets[index + 1])
            self.point_charge_forces = computed.point_charge_forces[pb:pe].copy()
            if computed.point_charge_potential_gradient is not None:
                self.point_charge_potential_gradient = \
                    computed.point_charge_potential_gradient[pb:pe].copy()
        self.dipole_moments = (computed.dipole_moments[index].copy()
                               if computed.dipole_moments is not None else None)
        self.scc_iterations = int(computed.scc_iterations[index])
        self.scc_converged = bool(computed.scc_converged[index])
        self.scc_status = int(computed.per_system_status[index])
        self._natoms = e - b

    _keys = ("energy", "energies", "forces", "gradient", "charges", "point_charge_forces",
             "dipole_moments", "scc_iterations", "scc_converged", "scc_status", "natoms")

    def get(self, attribute: str):
        if attribute == "energies":
            return np.asarray([self.energy])
        if attribute == "gradient":
            return -self.forces
        if attribute == "natoms":
            return self._natoms
        if attribute not in self._keys:
            raise XTBloomValueError(f"attribute {attribute!r} is not available in this result")
        return getattr(self, attribute)

    __getitem__ = get

    def dict(self):
        return {k: self.get(k) for k in self._keys}


class BatchResult:
    def __init__(self, computed: ComputedBatch):
        self._computed = computed
        self.energies = computed.energies
        self.scc_iterations = computed.scc_iterations
        self.scc_converged = computed.scc_converged
        self.per_system_status = computed.per_system_status

    def __len__(self):
        return len(self.energies)

    def __getitem__(self, index: int) -> Result:
        return Result(self._computed, index)

    def failed_indices(self) -> np.ndarray:
        return np.nonzero(self.per_system_status != library.STATUS_SUCCESS)[0]

    def raise_for_status(self):
        bad = self.failed_indices()
        if len(bad):
            raise XTBloomError(int(self.per_system_status[bad[0]]),
                               f"{len(bad)} system(s) failed, first index {int(bad[0])}")

    def get(self, attribute: str):
        c = self._computed
        table = {"energies": c.energies, "forces": c.forces, "gradient": -c.forces,
                 "charges": c.charges, "point_charge_forces": c.point_charge_forces,
                 "dipole_moments": c.dipole_moments, "scc_iterations": c.scc_iterations,
                 "scc_converged": c.scc_converged, "scc_status": c.per_system_status}
        if attribute not in table:
            raise XTBloomValueError(f"attribute {attribute!r} is not available in this result")
        return table[attribute]


def _flatten(structures: Sequence[Structure]):
    offs = np.zeros(len(structures) + 1, np.int64)
    offs[1:] = np.cumsum([len(s.numbers) for s in structures])
    numbers = np.concatenate([s.numbers for s in structures])
    positions = np.concatenate([s.positions for s in structures])
    charges = np.array([s.charge for s in structures])
    uhf = np.array([s.uhf for s in structures], np.int32)
    kw = {}
    if any(s.point_charges is not None for s in structures):
        poff = np.zeros(len(structures) + 1, np.int64)
        poff[1:] = np.cumsum([0 if s.point_charges is None else len(s.point_charges.charges)
                              for s in structures])
        pcs = [s.point_charges for s in structures if s.point_charges is not None]
        kw = dict(point_offsets=poff,
                  point_positions=np.concatenate([p.positions for p in pcs]),
                  point_values=np.concatenate([p.charges for p in pcs]),
                  point_gammas=np.concatenate([p.gammas for p in pcs]))
    if any(s.efield is not None for s in structures):
        kw["electric_field"] = [s.efield for s in structures]
    if any(s.spin_channels == 2 for s in structures):
        kw["spin_channels"] = np.array([s.spin_channels for s in structures], np.int32)
    if any(s.charge_response is not None for s in structures):
        # systems without an operator get zero b and A: the ABI takes one
        # batch-wide operator (reference python/xtbloom/interface.py:679-710)
        shifts, mats = [], []
        for s in structures:
            n = len(s.numbers)
            r = s.charge_response
            shifts.append(np.zeros(n) if r is None else r.shifts)
            mats.append(np.zeros(n * n) if r is None else r.matrix.ravel())
        roff = np.zeros(len(structures) + 1, np.int64)
        roff[1:] = np.cumsum([m.size for m in mats])
        kw.update(atomic_potential_shifts=np.concatenate(shifts), charge_response_offsets=roff,
                  charge_response_matrix=np.concatenate(mats))
    return offs, numbers, positions, charges, uhf, kw


class _Settings:
    def __init__(self, **kw):
        self.max_scc_iterations = kw.get("max_scc_iterations", 250)
        self.charge_tolerance = kw.get("charge_tolerance", 1.0e-6)
        self.energy_tolerance = kw.get("energy_tolerance", 1.0e-8)
        self.electronic_temperature = kw.get("electronic_temperature", 300.0)
        self.scc_mixer_history = kw.get("scc_mixer_history", library.DEFAULT_SCC_MIXER_HISTORY)
        self.scc_mixer_damping = kw.get("scc_mixer_damping", library.DEFAULT_SCC_MIXER_DAMPING)

    def set(self, attribute, value):
        aliases = {"max-iter": "max_scc_iterations", "temperature": "electronic_temperature"}
        attribute = aliases.get(attribute, attribute)
        if not hasattr(self, attribute):
            raise XTBloomValueError(f"unknown setting {attribute!r}")
        setattr(self, attribute, value)

    def kwargs(self):
        return dict(vars(self))


def _compute_auto_warm(context: Context, warm_start: bool, *args, **kw) -> ComputedBatch:
    """compute_arrays with the reference calculators' opt-in automatic warm start
    (reference python/xtbloom/interface.py:1274-1281, 1366-1388): WARM when the
    context holds a fully converged checkpoint, one FRESH retry when the native
    gate rejects it (changed identity), and the readiness flag mirrored from the
    call's statuses."""
    ready = warm_start and getattr(context, "_warm_ready", False)
    mode = library.SCC_START_WARM if ready else library.SCC_START_FRESH
    try:
        computed = compute_arrays(context, *args, scc_start_mode=mode, **kw)
    except XTBloomError as error:
        if mode != library.SCC_START_WARM or error.status != library.STATUS_INVALID_ARGUMENT:
            raise
        computed = compute_arrays(context, *args, scc_start_mode=library.SCC_START_FRESH, **kw)
    context._warm_ready = bool(np.all(computed.per_system_status == 0) and
                               np.all(computed.scc_converged))
    return computed


def _resolve_method(method: str) -> int:
    key = method.upper()
    if key not in _METHODS:
        raise XTBloomValueError(f"unknown method {method!r}")
    return _METHODS[key]


class Calculator(Structure):
    """Single-point GFN2-xTB calculator on the MI355X backend (tblite-like API)."""

    def __init__(self, method: str, numbers, positions, charge: float = 0.0, uhf: int | None = None,
                 point_charges: PointCharge | None = None, *, backend: str | int = "hip",
                 device_id: int | None = None, warm_start: bool = False, efield=None,
                 charge_response: ChargeResponse | None = None, spin_channels: int | None = None,
                 **settings):
        super().__init__(numbers, positions, charge, uhf, point_charges, efield, charge_response,
                         spin_channels)
        self._model = _resolve_method(method)
        self._method = method
        self._context = Context(backend, device_id)
        self._settings = _Settings(**settings)
        # warm_start=True seeds each SCC from the previous fully converged call
        # with the same topology and settings (native SCC_START_WARM)
        self._warm_start = bool(warm_start)

    @property
    def backend(self) -> int:
        return self._context.backend

    @property
    def method(self) -> str:
        return self._method

    def set(self, attribute: str, value):
        self._settings.set(attribute, value)

    def singlepoint(self) -> Result:
        flags = library.COMPUTE_ENERGY | library.COMPUTE_FORCES | library.COMPUTE_ATOMIC_CHARGES
        if self.point_charges is not None:
            flags |= library.COMPUTE_POINT_CHARGE_FORCES
        if self.efield is not None:
            # a field run publishes the molecular dipole (reference interface.py:1926-1927)
            flags |= library.COMPUTE_DIPOLE_MOMENTS
        offs, numbers, positions, charges, uhf, kw = _flatten([self])
        computed = _compute_auto_warm(self._context, self._warm_start, offs, numbers, positions,
                                      charges, uhf, **kw, model=self._model, flags=flags,
                                      extension=True, **self._settings.kwargs())
        BatchResult(computed).raise_for_status()
        return Result(computed, 0)

    def close(self):
        self._context.close()


class BatchCalculator:
    """Batched GFN2-xTB over many structures in one native call."""

    def __init__(self, structures: Sequence[Structure], method: str = "GFN2-xTB", *,
                 backend: str | int = "hip", device_id: int | None = None,
                 warm_start: bool = False, **settings):
        if not structures:
            raise XTBloomValueError("a batch needs at least one structure")
        self._structures = list(structures)
        self._model = _resolve_method(method)
        self._context = Context(backend, device_id)
        self._settings = _Settings(**settings)
        self._warm_start = bool(warm_start)

    @property
    def backend(self) -> int:
        return self._context.backend

    def set(self, attribute, value):
        self._settings.set(attribute, value)

    def compute(self, *, flags=None) -> BatchResult:
        offs, numbers, positions, charges, uhf, kw = _flatten(self._structures)
        if flags is None:
            flags = library.COMPUTE_ENERGY | library.COMPUTE_FORCES | library.COMPUTE_ATOMIC_CHARGES
            if "point_offsets" in kw:
                flags |= library.COMPUTE_POINT_CHARGE_FORCES
            if "electric_field" in kw:
                flags |= library.COMPUTE_DIPOLE_MOMENTS
        computed = _compute_auto_warm(self._context, self._warm_start, offs, numbers, positions,
                                      charges, uhf, **kw, model=self._model, flags=flags,
                                      **self._settings.kwargs())
        return BatchResult(computed)

    def close(self):
        self._context.close()
