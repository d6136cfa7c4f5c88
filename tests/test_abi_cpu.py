"""C-ABI checks that need no GPU: the library loads, exports every declared
entry point, keeps the reference struct images, and refuses misuse without
touching a device."""
import ctypes
import os
import re

import pytest

from xtbloom_amd import library

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for header in ("include/xtbloom/xtbloom.h", "include/xtbloom_amd/extensions.h"):
        text = open(os.path.join(ROOT, header)).read()
        names.update(re.findall(r"XTBLOOM_API[^;(]*?\b(xtbloom_\w+)\s*\(", text, re.S))
    return names


def test_library_exports_every_declared_symbol():
    lib = library.load()
    names = declared_symbols()
    assert len(names) >= 33
    for name in sorted(names):
        assert hasattr(lib, name), name
    assert set(library.EXPORTED_SYMBOLS) <= names


def test_struct_images_match_reference_lp64():
    assert ctypes.sizeof(library.ContextOptions) == 32
    assert ctypes.sizeof(library.Batch) == 456
    assert library.Batch.spin_channels.offset == 328
    assert library.Batch.total_interactions.offset == 352
    assert library.Batch.cell_matrices.offset == 408
    assert ctypes.sizeof(library.ComputeOptions) == 80
    assert library.ComputeOptions.scc_start_mode.offset == 48
    assert library.ComputeOptions.scc_mixer_damping.offset == 64
    assert ctypes.sizeof(library.BatchResult) == 280
    assert library.BatchResult.dipole_moments.offset == 184


def test_option_defaults_follow_reference():
    lib = library.load()
    o = library.ComputeOptions()
    assert lib.xtbloom_compute_options_init(ctypes.byref(o), ctypes.sizeof(o)) == 0
    assert o.model == library.MODEL_GFN2_XTB
    assert o.flags == library.COMPUTE_ENERGY | library.COMPUTE_FORCES
    assert o.max_scc_iterations == 250
    assert o.charge_tolerance == 1e-6 and o.energy_tolerance == 1e-8
    assert o.electronic_temperature == 300.0 * library.KELVIN_TO_HARTREE
    assert o.scc_start_mode == library.SCC_START_FRESH
    assert o.scc_mixer_history == 8 and o.scc_mixer_damping == 0.4
    # too-small struct is rejected and reported
    assert lib.xtbloom_compute_options_init(ctypes.byref(o), 8) == library.STATUS_INVALID_ARGUMENT
    assert "smaller than ABI v1" in library.last_error()


def test_cpu_and_cuda_backends_are_unavailable():
    lib = library.load()
    opts = library.ContextOptions()
    lib.xtbloom_context_options_init(ctypes.byref(opts), ctypes.sizeof(opts))
    for backend in (library.BACKEND_CPU, library.BACKEND_CUDA):
        opts.backend = backend
        handle = ctypes.c_void_p()
        st = lib.xtbloom_context_create(ctypes.byref(opts), ctypes.byref(handle))
        assert st == library.STATUS_BACKEND_UNAVAILABLE
        assert not handle.value


def test_compute_rejects_null_context():
    lib = library.load()
    batch, opts, res = library.Batch(), library.ComputeOptions(), library.BatchResult()
    lib.xtbloom_batch_init(ctypes.byref(batch), ctypes.sizeof(batch))
    lib.xtbloom_compute_options_init(ctypes.byref(opts), ctypes.sizeof(opts))
    lib.xtbloom_batch_result_init(ctypes.byref(res), ctypes.sizeof(res))
    st = lib.xtbloom_compute(None, ctypes.byref(batch), ctypes.byref(opts), ctypes.byref(res))
    assert st == library.STATUS_INVALID_ARGUMENT


def test_status_strings():
    assert library.status_string(0) == "success"
    assert library.status_string(8) == "eigensolver failed"
    assert library.status_string(99) == "unknown status"


def test_packed_matrix_elements_closed_form():
    import numpy as np
    from xtbloom_amd import workloads
    lib = library.load()
    b = workloads.make("ragged512", seed=0, batch=20)
    batch = library.Batch()
    lib.xtbloom_batch_init(ctypes.byref(batch), ctypes.sizeof(batch))
    offs = np.ascontiguousarray(b.atom_offsets)
    Z = np.ascontiguousarray(b.atomic_numbers)
    batch.batch_size = b.batch_size
    batch.atom_offsets.data, batch.atom_offsets.size_bytes = offs.ctypes.data, offs.nbytes
    batch.atomic_numbers.data, batch.atomic_numbers.size_bytes = Z.ctypes.data, Z.nbytes
    expect = 0
    for i in range(b.batch_size):
        z = b.system(i)["Z"]
        nao = sum(1 if x == 1 else 4 for x in z)  # H: 1s; C/N/O: 2s + 2p
        expect += nao * nao
    assert lib.xtbloom_amd_packed_matrix_elements(ctypes.byref(batch)) == expect
