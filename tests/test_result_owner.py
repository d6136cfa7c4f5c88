"""Host result arenas and their DLPack export through the C ABI.

The cases follow the reference's tests/result_owner_test.cpp: options
defaults and short-struct rejection, reference counting, the managed-tensor
layouts of DLPack 1.0 (versioned and legacy) read field by field, independent
repeated exports, exports outliving the creator's reference, and the
transactional failures (out-of-range view, unsupported or undocumented dtype,
unaligned offset, negative extent, unknown version, NULL owner).  Host arenas
need no GPU; the device arena is covered by the GPU tests.
"""
import ctypes
import gc

import numpy as np
import pytest

from xtbloom_amd import dlpack, library

lib = library.load()
vp = ctypes.c_void_p


class DLTensor(ctypes.Structure):
    _fields_ = [("data", vp), ("device_type", ctypes.c_int32), ("device_id", ctypes.c_int32),
                ("ndim", ctypes.c_int32), ("code", ctypes.c_uint8), ("bits", ctypes.c_uint8),
                ("lanes", ctypes.c_uint16), ("shape", ctypes.POINTER(ctypes.c_int64)),
                ("strides", ctypes.POINTER(ctypes.c_int64)), ("byte_offset", ctypes.c_uint64)]


DELETER = ctypes.CFUNCTYPE(None, vp)


class Versioned(ctypes.Structure):
    _fields_ = [("major", ctypes.c_uint32), ("minor", ctypes.c_uint32), ("manager_ctx", vp),
                ("deleter", vp), ("flags", ctypes.c_uint64), ("dl_tensor", DLTensor)]


class Legacy(ctypes.Structure):
    _fields_ = [("dl_tensor", DLTensor), ("manager_ctx", vp), ("deleter", vp)]


assert ctypes.sizeof(DLTensor) == 48 and ctypes.sizeof(Legacy) == 64 and ctypes.sizeof(Versioned) == 80


def options(size, space=library.MEMORY_HOST, device=-1):
    o = library.ResultOwnerOptions()
    assert lib.xtbloom_result_owner_options_init(ctypes.byref(o), ctypes.sizeof(o)) == 0
    o.memory_space, o.device_id, o.size_bytes = space, device, size
    return o


def create(size):
    h = vp()
    assert lib.xtbloom_result_owner_create(ctypes.byref(options(size)), ctypes.byref(h)) == 0 and h.value
    return h


def view(offset, code, bits, shape):
    ext = (ctypes.c_int64 * max(1, len(shape)))(*shape)
    v = library.DLPackView(ctypes.sizeof(library.DLPackView), library.API_VERSION, offset, code, bits, 1,
                           len(shape), 0, ext)
    v._keep = ext
    return v


def export(owner, v, version):
    out = vp(0x1234)
    st = lib.xtbloom_result_owner_export_dltensor(owner, ctypes.byref(v), version, ctypes.byref(out))
    return st, out.value


def delete(managed, cls):
    DELETER(cls.from_address(managed).deleter)(managed)


def test_options_defaults_and_short_struct():
    o = library.ResultOwnerOptions()
    assert lib.xtbloom_result_owner_options_init(ctypes.byref(o), ctypes.sizeof(o)) == 0
    assert o.memory_space == library.MEMORY_HOST and o.device_id == -1
    assert lib.xtbloom_result_owner_options_init(ctypes.byref(o), 27) == library.STATUS_INVALID_ARGUMENT
    v = library.DLPackView()
    assert lib.xtbloom_result_owner_export_dltensor(None, ctypes.byref(v), 1, None) == \
        library.STATUS_INVALID_ARGUMENT


def test_host_arena_versioned_and_legacy_exports():
    owner = create(4096)
    buf = library.Buffer()
    assert lib.xtbloom_result_owner_buffer(owner, ctypes.byref(buf)) == 0
    assert buf.data and buf.size_bytes == 4096 and buf.memory_space == library.MEMORY_HOST
    assert buf.reserved == 0
    v = view(64, 2, 64, (8, 8))
    st, m1 = export(owner, v, 1)
    assert st == 0 and m1
    t = Versioned.from_address(m1)
    assert (t.major, t.minor, t.flags) == (1, 0, 0)
    d = t.dl_tensor
    assert d.ndim == 2 and d.device_type == 1 and d.device_id == 0
    assert (d.code, d.bits, d.lanes) == (2, 64, 1)
    assert d.byte_offset == 0 and d.data == buf.data + 64  # the offset is folded into data
    assert d.shape[0] == 8 and d.shape[1] == 8 and not d.strides
    st, m2 = export(owner, v, 1)
    assert st == 0 and m2 and m2 != m1
    st, m3 = export(owner, v, 0)
    assert st == 0 and m3
    legacy = Legacy.from_address(m3)
    assert legacy.dl_tensor.ndim == 2 and legacy.dl_tensor.device_type == 1 and legacy.manager_ctx
    # the exports keep the arena alive past the creator's reference
    lib.xtbloom_result_owner_release(owner)
    (ctypes.c_double * 64).from_address(d.data)[63] = 2.5
    delete(m1, Versioned)
    delete(m2, Versioned)
    delete(m3, Legacy)


def test_reference_counting_and_failures():
    h = create(512)
    lib.xtbloom_result_owner_retain(h)
    lib.xtbloom_result_owner_retain(h)
    lib.xtbloom_result_owner_release(h)
    lib.xtbloom_result_owner_release(h)
    lib.xtbloom_result_owner_release(h)  # freed here
    before = library.last_error()
    lib.xtbloom_result_owner_release(None)
    assert library.last_error() == before
    out = vp()
    for size, space, device, reserved in ((0, library.MEMORY_HOST, -1, 0), (512, library.MEMORY_HOST, -1, 1),
                                          (512, library.MEMORY_HOST, 0, 0)):
        o = options(size, space, device)
        o.reserved = reserved
        assert lib.xtbloom_result_owner_create(ctypes.byref(o), ctypes.byref(out)) == \
            library.STATUS_INVALID_ARGUMENT and not out.value
    owner = create(512)
    bad = [view(0, 2, 64, (8, 64)),  # 4096 bytes > 512
           view(0, 3, 64, (16,)),  # unsupported code
           view(4, 2, 64, (16,)),  # unaligned offset
           view(0, 2, 64, (-1,))]  # negative extent
    bad += [view(0, c, b, (16,)) for c, b in ((1, 16), (1, 32), (1, 64), (4, 16), (6, 8))]
    for v in bad:
        assert export(owner, v, 1) == (library.STATUS_INVALID_ARGUMENT, None)
    assert export(owner, view(0, 2, 64, (8, 8)), 7) == (library.STATUS_INVALID_ARGUMENT, None)
    # an empty view is valid: NULL data and the copied extents
    st, m = export(owner, view(0, 2, 64, (0, 3)), 1)
    assert st == 0 and m
    t = Versioned.from_address(m)
    assert not t.dl_tensor.data and t.dl_tensor.shape[0] == 0 and t.dl_tensor.shape[1] == 3
    delete(m, Versioned)
    buf = library.Buffer()
    assert lib.xtbloom_result_owner_buffer(owner, ctypes.byref(buf)) == 0 and buf.size_bytes == 512
    lib.xtbloom_result_owner_release(owner)


@pytest.mark.parametrize("dtype", [np.float64, np.float32, np.int32, np.int64, np.uint8, np.int8])
def test_numpy_imports_host_arena(dtype):
    owner = dlpack.ResultOwner(4096)
    n = (4096 - 64) // np.dtype(dtype).itemsize
    src = (np.arange(n) % 100).astype(dtype)
    ctypes.memmove(owner.data + 64, src.ctypes.data, src.nbytes)
    got = np.from_dlpack(owner.tensor(64, dtype, (n,)))
    assert got.dtype == np.dtype(dtype) and np.array_equal(got, src)
    owner.close()  # the imported array holds its own reference
    assert np.array_equal(got, src)
    del got
    gc.collect()
