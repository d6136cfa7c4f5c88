"""Host-side pieces of the Python interface that need no GPU: electric-field
normalization and attachment packing (reference python/xtbloom/interface.py:
273-289, 1172-1184), ChargeResponse validation (interface.py:231-267), the
open-shell spin-channel default (interface.py:158-162), and the batch
descriptors these produce."""
import ctypes

import numpy as np
import pytest

from xtbloom_amd import interface, library

lib = library.load()


def test_efield_normalization():
    assert interface._normalize_efield(None) is None
    assert interface._normalize_efield([0.0, 0.0, 0.0]) is None  # zero field: no attachment
    f = interface._normalize_efield([0.001, 0.0, -0.002])
    assert f.tolist() == [0.001, 0.0, -0.002] and not f.flags.writeable
    for bad in ([0.1, 0.2], [[0.1, 0.2, 0.3]], [0.1, np.nan, 0.0], [np.inf, 0.0, 0.0]):
        with pytest.raises(interface.XTBloomValueError):
            interface._normalize_efield(bad)


def test_field_attachment_packing():
    count, desc, payload = interface._field_attachments(4, [None, (1.0, 2.0, 3.0), None, (0.0, 0.0, 0.0)])
    assert count == 2
    assert [desc[k].system_index for k in range(2)] == [1, 3]
    for k in range(2):
        assert desc[k].type == library.INTERACTION_ELECTRIC_FIELD and desc[k].flags == 0
        assert desc[k].payload_offset == 32 * k and desc[k].payload_size == 32
    raw = payload.tobytes()
    head = np.frombuffer(raw, np.int32)
    assert head[0] == 1 and head[1] == 0 and head[8] == 1 and head[9] == 0
    vals = np.frombuffer(raw, np.float64)
    assert vals[1:4].tolist() == [1.0, 2.0, 3.0] and vals[5:8].tolist() == [0.0, 0.0, 0.0]
    with pytest.raises(interface.XTBloomValueError):
        interface._field_attachments(2, [None])
    with pytest.raises(interface.XTBloomValueError):
        interface._field_attachments(1, [(1.0, 2.0)])


def _descriptors(**kw):
    offs = np.array([0, 3, 5], np.int64)
    numbers = np.array([8, 1, 1, 1, 1], np.int32)
    pos = np.arange(15, dtype=float).reshape(5, 3)
    return interface._descriptors(lib, offs, numbers, pos, None, None, None, None, None, None,
                                  library.MODEL_GFN2_XTB, interface._DEFAULT_FLAGS, 250, 1e-6, 1e-8, 300.0,
                                  8, 0.4, library.SCC_START_FRESH, **kw)


def test_descriptors_bind_fields_operator_and_spin():
    batch, opts, keep, (B, N, P) = _descriptors()
    assert (B, N, P) == (2, 5, 0)
    assert batch.total_interactions == 0 and not batch.interaction_descriptors.data
    assert not batch.spin_channels.data and batch.total_charge_response_elements == 0
    batch, _, keep, _ = _descriptors(electric_field=[(0.1, 0.0, 0.0), None],
                                     atomic_potential_shifts=np.zeros(5),
                                     charge_response_offsets=np.array([0, 9, 13]),
                                     charge_response_matrix=np.zeros(13),
                                     spin_channels=[2, 1])
    assert batch.total_interactions == 1
    assert batch.interaction_descriptors.size_bytes >= ctypes.sizeof(library.Interaction)
    assert batch.total_charge_response_elements == 13
    assert batch.atomic_potential_shifts.size_bytes == 40
    assert batch.spin_channels.size_bytes == 8
    assert np.frombuffer((ctypes.c_char * 8).from_address(batch.spin_channels.data).raw, np.int32).tolist() == [2, 1]


def test_charge_response_validation():
    r = interface.ChargeResponse([0.1, 0.2], [[1.0, 0.5], [0.5, 2.0]])
    assert r.shifts.dtype == np.float64 and r.matrix.shape == (2, 2)
    for shifts, matrix in (([[0.1]], [[1.0]]), ([0.1, 0.2], [[1.0, 0.5]]), ([0.1], [[1.0, 0.0], [0.0, 1.0]]),
                           ([np.nan], [[1.0]]), ([0.1], [[np.inf]])):
        with pytest.raises(interface.XTBloomValueError):
            interface.ChargeResponse(shifts, matrix)


def test_structure_spin_default_and_flatten():
    water = interface.Structure([8, 1, 1], np.zeros((3, 3)) + np.arange(3)[:, None])
    oh = interface.Structure([8, 1], [[0, 0, 0], [1.8, 0, 0]], uhf=1)
    oh_r = interface.Structure([8, 1], [[0, 0, 0], [1.8, 0, 0]], uhf=1, spin_channels=1)
    assert water.spin_channels == 1 and oh.spin_channels == 2 and oh_r.spin_channels == 1
    with pytest.raises(interface.XTBloomValueError):
        interface.Structure([8, 1], [[0, 0, 0], [1.8, 0, 0]], uhf=1, spin_channels=0)
    *_, kw = interface._flatten([water, oh_r])
    assert "spin_channels" not in kw  # every member restricted: no buffer
    *_, kw = interface._flatten([water, oh])
    assert kw["spin_channels"].tolist() == [1, 2]
    resp = interface.ChargeResponse([0.1, 0.2], [[1.0, 0.5], [0.5, 2.0]])
    f = interface.Structure([8, 1], [[0, 0, 0], [1.8, 0, 0]], uhf=1, efield=[0.0, 0.0, 0.01],
                            charge_response=resp)
    offs, numbers, positions, charges, uhf, kw = interface._flatten([water, f])
    assert kw["electric_field"][0] is None and kw["electric_field"][1].tolist() == [0.0, 0.0, 0.01]
    assert kw["charge_response_offsets"].tolist() == [0, 9, 13]
    assert kw["atomic_potential_shifts"].tolist() == [0.0, 0.0, 0.0, 0.1, 0.2]
    assert kw["charge_response_matrix"][9:].tolist() == [1.0, 0.5, 0.5, 2.0]
    with pytest.raises(interface.XTBloomValueError):
        interface.Structure([8, 1, 1], np.zeros((3, 3)), charge_response=resp)
