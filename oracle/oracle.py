"""ctypes binding of the C oracle (``liboracle_gfn2.so``).

TEST INFRASTRUCTURE ONLY.  Imported by tests/, ``__graft_entry__.smoke()`` and
the cpu_baseline leg of ``bench.py`` as the checker / CPU baseline; the
product package never imports it.
"""
from __future__ import annotations

import ctypes
import glob
import os
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle_gfn2.so")

_c_double_p = ctypes.POINTER(ctypes.c_double)
_c_int32_p = ctypes.POINTER(ctypes.c_int32)
_c_int64_p = ctypes.POINTER(ctypes.c_int64)


class Options(ctypes.Structure):
    _fields_ = [("max_iterations", ctypes.c_int32), ("history", ctypes.c_int32),
                ("charge_tolerance", ctypes.c_double), ("energy_tolerance", ctypes.c_double),
                ("kt", ctypes.c_double), ("damping", ctypes.c_double),
                ("efield", ctypes.c_double * 3), ("potential_shifts", _c_double_p),
                ("charge_response", _c_double_p), ("spin_channels", ctypes.c_int32),
                ("reserved", ctypes.c_int32)]


class Trace(ctypes.Structure):
    _fields_ = [("capacity", ctypes.c_int32), ("recorded", ctypes.c_int32)] + [
        (name, _c_double_p) for name in (
            "free_energy", "residual_rms", "overlap", "h0", "dipole", "quadrupole", "density",
            "shell_charges", "dipoles", "quadrupoles", "coordination", "components")]


def lapack_path() -> str:
    import scipy
    hits = sorted(glob.glob(os.path.join(os.path.dirname(scipy.__file__), "..", "scipy.libs",
                                         "libscipy_openblas*.so")))
    if not hits:
        raise RuntimeError("scipy-openblas LP64 provider not found")
    return os.path.abspath(hits[0])


_LIB = None


def build() -> str:
    """Compile the oracle with gcc (idempotent; called by __graft_entry__.build)."""
    import subprocess
    src = os.path.join(HERE, "gfn2_oracle.c")
    cmd = ["gcc", "-O2", "-std=c11", "-D_GNU_SOURCE", "-fPIC", "-shared", "-o", LIB_PATH, src,
           "-ldl", "-lpthread", "-lm"]
    subprocess.check_call(cmd)
    return LIB_PATH


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        L.oracle_init.argtypes = [ctypes.c_char_p]
        if L.oracle_init(lapack_path().encode()) != 0:
            raise RuntimeError("oracle could not load the LAPACK provider")
        L.oracle_singlepoint.argtypes = [
            ctypes.c_int32, _c_int32_p, _c_double_p, ctypes.c_double, ctypes.c_int32,
            ctypes.c_int32, _c_double_p, _c_double_p, _c_double_p, ctypes.POINTER(Options),
            _c_double_p, _c_double_p, _c_double_p, _c_double_p, _c_double_p, _c_int32_p,
            ctypes.POINTER(Trace)]
        L.oracle_singlepoint_state.argtypes = L.oracle_singlepoint.argtypes + [
            _c_double_p, _c_double_p]
        L.oracle_batch.argtypes = [
            ctypes.c_int64, _c_int64_p, _c_int32_p, _c_double_p, _c_double_p, _c_int32_p,
            _c_int64_p, _c_double_p, _c_double_p, _c_double_p, ctypes.POINTER(Options),
            ctypes.c_int32, _c_double_p, _c_double_p, _c_double_p, _c_double_p, _c_int32_p,
            _c_int32_p, _c_double_p, _c_double_p, _c_double_p, _c_int32_p]
        L.oracle_scc_replay.argtypes = [
            ctypes.c_int32, _c_int32_p, _c_double_p, ctypes.c_double, ctypes.c_int32,
            ctypes.c_int32, _c_double_p, _c_double_p, _c_double_p, ctypes.c_double, _c_double_p,
            _c_double_p, _c_double_p, _c_double_p, _c_double_p]
        L.oracle_scc_replay_channels.argtypes = [
            ctypes.c_int32, _c_int32_p, _c_double_p, ctypes.c_double, ctypes.c_int32, ctypes.c_int32,
            ctypes.c_int32, _c_double_p, _c_double_p, _c_double_p, ctypes.c_double, _c_double_p,
            _c_double_p, _c_double_p, _c_double_p, _c_double_p, _c_double_p, _c_double_p]
        L.oracle_broyden.argtypes = [ctypes.c_int32, ctypes.c_int32, ctypes.c_double,
                                     _c_double_p, ctypes.c_int32, _c_double_p, _c_double_p]
        L.oracle_basis_counts.argtypes = [ctypes.c_int32, _c_int32_p, _c_int32_p, _c_int32_p,
                                          _c_double_p]
        L.oracle_coordination.argtypes = [ctypes.c_int32, _c_int32_p, _c_double_p, _c_double_p]
        L.oracle_integrals.argtypes = [ctypes.c_int32, _c_int32_p, _c_double_p, _c_double_p,
                                       _c_double_p, _c_double_p]
        L.oracle_h0.argtypes = [ctypes.c_int32, _c_int32_p, _c_double_p, _c_double_p, _c_double_p]
        L.oracle_repulsion.argtypes = [ctypes.c_int32, _c_int32_p, _c_double_p, _c_double_p]
        L.oracle_repulsion.restype = ctypes.c_double
        L.oracle_d4_two_body.argtypes = [ctypes.c_int32, _c_int32_p, _c_double_p, _c_double_p,
                                         _c_double_p, _c_double_p, _c_double_p]
        L.oracle_d4_atm.argtypes = [ctypes.c_int32, _c_int32_p, _c_double_p, _c_double_p,
                                    _c_double_p]
        L.oracle_fermi.argtypes = [ctypes.c_int32, _c_double_p, ctypes.c_double, ctypes.c_double,
                                   _c_double_p, _c_double_p, _c_double_p]
        _LIB = L
    return _LIB


def _p(a, t=_c_double_p):
    return None if a is None else a.ctypes.data_as(t)


def options(max_iterations=250, charge_tolerance=1e-6, energy_tolerance=1e-8,
            kt=300.0 * 3.166808578545117e-6, history=8, damping=0.4,
            efield=(0.0, 0.0, 0.0), shifts=None, response=None, spin_channels=1) -> Options:
    """``shifts`` (b, per atom) and ``response`` (A, nat x nat symmetric) add
    the external charge operator v = b + A q, E = q.b + q.A.q/2."""
    o = Options(max_iterations, history, charge_tolerance, energy_tolerance, kt, damping,
                (ctypes.c_double * 3)(*efield))
    o.spin_channels = spin_channels
    keep = []
    if shifts is not None:
        keep.append(np.ascontiguousarray(shifts, dtype=np.float64))
        o.potential_shifts = _p(keep[-1])
    if response is not None:
        keep.append(np.ascontiguousarray(response, dtype=np.float64))
        o.charge_response = _p(keep[-1])
    o._keep = keep  # the arrays live as long as the options
    return o


@dataclass
class Result:
    status: int
    energy: float
    gradient: np.ndarray
    charges: np.ndarray
    iterations: int
    pc_gradient: np.ndarray | None = None
    dedq: np.ndarray | None = None
    trace: dict | None = None
    state: np.ndarray | None = None


def basis_counts(Z):
    Z = np.ascontiguousarray(Z, dtype=np.int32)
    nsh, nao, nel = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_double()
    st = lib().oracle_basis_counts(len(Z), _p(Z, _c_int32_p), ctypes.byref(nsh),
                                   ctypes.byref(nao), ctypes.byref(nel))
    if st != 0:
        raise ValueError("unsupported element")
    return nsh.value, nao.value, nel.value


def singlepoint(Z, xyz, charge=0.0, uhf=0, pc_xyz=None, pc_q=None, pc_gamma=None, opts=None,
                trace=False, trace_capacity=512, warm_state=None) -> Result:
    """One single point.  ``warm_state`` (the ``state`` of an earlier converged
    result of the same system) starts the SCC from that state instead of the
    FRESH SAD guess; ``Result.state`` is always the final SCC multipole state."""
    Z = np.ascontiguousarray(Z, dtype=np.int32)
    xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
    nat = len(Z)
    npc = 0 if pc_q is None else len(pc_q)
    if npc:
        pc_xyz = np.ascontiguousarray(pc_xyz, dtype=np.float64).reshape(-1, 3)
        pc_q = np.ascontiguousarray(pc_q, dtype=np.float64)
        pc_gamma = np.ascontiguousarray(pc_gamma, dtype=np.float64)
    e = ctypes.c_double(np.nan)
    g = np.zeros((nat, 3))
    q = np.zeros(nat)
    pg = np.zeros((npc, 3)) if npc else None
    dq = np.zeros(npc) if npc else None
    it = ctypes.c_int32(0)
    tr = None
    bufs = {}
    if trace:
        nsh, nao, _ = basis_counts(Z)
        nn = nao * nao
        shapes = dict(free_energy=(trace_capacity,), residual_rms=(trace_capacity,),
                      overlap=(nao, nao), h0=(nao, nao), dipole=(3, nao, nao),
                      quadrupole=(6, nao, nao), density=(nao, nao), shell_charges=(nsh,),
                      dipoles=(nat, 3), quadrupoles=(nat, 6), coordination=(nat,),
                      components=(8,))
        bufs = {k: np.zeros(s) for k, s in shapes.items()}
        tr = Trace(trace_capacity, 0, *[_p(bufs[k]) for k in (
            "free_energy", "residual_rms", "overlap", "h0", "dipole", "quadrupole", "density",
            "shell_charges", "dipoles", "quadrupoles", "coordination", "components")])
    o = opts or options()
    nsh, _, _ = basis_counts(Z)
    # the state holds every spin channel of [qsh, dip(3N), quad(6N)]
    state = np.zeros(int(o.spin_channels) * (nsh + 9 * nat))
    if warm_state is not None:
        warm_state = np.ascontiguousarray(warm_state, dtype=np.float64)
        if warm_state.shape != state.shape:
            raise ValueError("warm_state does not belong to this system")
    st = lib().oracle_singlepoint_state(
        nat, _p(Z, _c_int32_p), _p(xyz), float(charge), int(uhf), npc, _p(pc_xyz), _p(pc_q),
        _p(pc_gamma), ctypes.byref(o), ctypes.byref(e), _p(g), _p(q), _p(pg), _p(dq),
        ctypes.byref(it), ctypes.byref(tr) if tr is not None else None, _p(warm_state), _p(state))
    if trace:
        bufs["recorded"] = tr.recorded
        bufs["free_energy"] = bufs["free_energy"][:tr.recorded]
        bufs["residual_rms"] = bufs["residual_rms"][:tr.recorded]
    return Result(st, e.value, g, q, it.value, pg, dq, bufs if trace else None, state)


def broyden(initial, residuals, memory=8, damping=0.4):
    initial = np.ascontiguousarray(initial, dtype=np.float64)
    residuals = np.ascontiguousarray(residuals, dtype=np.float64)
    out = np.zeros_like(residuals)
    st = lib().oracle_broyden(len(initial), memory, damping, _p(initial), len(residuals),
                              _p(residuals), _p(out))
    if st != 0:
        raise RuntimeError("mixer failed")
    return out


def scc_replay(Z, xyz, mixed, charge=0.0, uhf=0, pc_xyz=None, pc_q=None, pc_gamma=None,
               kt=300.0 * 3.166808578545117e-6):
    """One SCC step from the mixed state [qsh, dip(3N), quad(6N)]; returns (raw, H, P, F)."""
    Z = np.ascontiguousarray(Z, dtype=np.int32)
    xyz = np.ascontiguousarray(xyz, dtype=np.float64)
    mixed = np.ascontiguousarray(mixed, dtype=np.float64)
    nsh, nao, _ = basis_counts(Z)
    npc = 0 if pc_q is None else len(pc_q)
    if npc:
        pc_xyz = np.ascontiguousarray(pc_xyz, dtype=np.float64)
        pc_q = np.ascontiguousarray(pc_q, dtype=np.float64)
        pc_gamma = np.ascontiguousarray(pc_gamma, dtype=np.float64)
    raw = np.zeros_like(mixed)
    H = np.zeros((nao, nao))
    P = np.zeros((nao, nao))
    F = ctypes.c_double()
    st = lib().oracle_scc_replay(len(Z), _p(Z, _c_int32_p), _p(xyz), float(charge), int(uhf), npc,
                                 _p(pc_xyz), _p(pc_q), _p(pc_gamma), float(kt), _p(mixed), _p(raw),
                                 _p(H), _p(P), ctypes.byref(F))
    if st != 0:
        raise RuntimeError(f"oracle replay failed with status {st}")
    return raw, H, P, F.value


def scc_replay_channels(Z, xyz, mixed, charge=0.0, uhf=0, spin_channels=1,
                        kt=300.0 * 3.166808578545117e-6):
    """One SCC step with 1 or 2 spin channels from a state in the reference's
    packing ([qsh, dip, quad], each field's channels in turn); returns raw,
    H [nch, nao, nao], P [nch, nao, nao] (alpha, beta; restricted: total),
    eigenvalues [nch, nao], occupations [2, nao] and F."""
    Z = np.ascontiguousarray(Z, dtype=np.int32)
    xyz = np.ascontiguousarray(xyz, dtype=np.float64)
    mixed = np.ascontiguousarray(mixed, dtype=np.float64)
    _, nao, _ = basis_counts(Z)
    nch = int(spin_channels)
    raw = np.zeros_like(mixed)
    H = np.zeros((nch, nao, nao))
    P = np.zeros((nch, nao, nao))
    eps = np.zeros((nch, nao))
    occ = np.zeros((2, nao))
    F = ctypes.c_double()
    st = lib().oracle_scc_replay_channels(len(Z), _p(Z, _c_int32_p), _p(xyz), float(charge), int(uhf), nch,
                                          0, None, None, None, float(kt), _p(mixed), _p(raw), _p(H), _p(P),
                                          _p(eps), _p(occ), ctypes.byref(F))
    if st != 0:
        raise RuntimeError(f"oracle replay failed with status {st}")
    return raw, H, P, eps, occ, F.value


def batch(atom_offsets, Z, xyz, charges, uhf, point_offsets=None, pc_xyz=None, pc_q=None,
          pc_gamma=None, opts=None, threads=0, want_point_gradient=False, efield=None,
          shifts=None, response=None, spin_channels=None):
    atom_offsets = np.ascontiguousarray(atom_offsets, dtype=np.int64)
    B = len(atom_offsets) - 1
    Z = np.ascontiguousarray(Z, dtype=np.int32)
    xyz = np.ascontiguousarray(xyz, dtype=np.float64)
    charges = np.ascontiguousarray(charges, dtype=np.float64)
    uhf = np.ascontiguousarray(uhf, dtype=np.int32)
    e = np.zeros(B)
    g = np.zeros((len(Z), 3))
    q = np.zeros(len(Z))
    it = np.zeros(B, dtype=np.int32)
    st = np.zeros(B, dtype=np.int32)
    pg = None
    if point_offsets is not None:
        point_offsets = np.ascontiguousarray(point_offsets, dtype=np.int64)
        pc_xyz = np.ascontiguousarray(pc_xyz, dtype=np.float64)
        pc_q = np.ascontiguousarray(pc_q, dtype=np.float64)
        pc_gamma = np.ascontiguousarray(pc_gamma, dtype=np.float64)
        if want_point_gradient:
            pg = np.zeros((len(pc_q), 3))
    o = opts or options()
    if efield is not None:  # [B, 3] uniform fields (zero rows: none)
        efield = np.ascontiguousarray(efield, dtype=np.float64).reshape(B, 3)
    if shifts is not None:  # [N] b per atom
        shifts = np.ascontiguousarray(shifts, dtype=np.float64)
    if response is not None:  # A of each system, nat x nat, packed in system order
        response = np.ascontiguousarray(response, dtype=np.float64)
    spin = None if spin_channels is None else np.ascontiguousarray(spin_channels, dtype=np.int32)
    lib().oracle_batch(B, _p(atom_offsets, _c_int64_p), _p(Z, _c_int32_p), _p(xyz), _p(charges),
                       _p(uhf, _c_int32_p), _p(point_offsets, _c_int64_p), _p(pc_xyz), _p(pc_q),
                       _p(pc_gamma), ctypes.byref(o), int(threads), _p(e), _p(g), _p(q), _p(pg),
                       _p(it, _c_int32_p), _p(st, _c_int32_p), _p(efield), _p(shifts),
                       _p(response), _p(spin, _c_int32_p))
    return dict(energies=e, gradient=g, charges=q, iterations=it, status=st, pc_gradient=pg)
