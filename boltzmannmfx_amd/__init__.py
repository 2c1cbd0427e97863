"""MI355X particle-mechanics path of BMX (BoltzmannMFX), Python bindings.

The product is the C-ABI library ``libbmx_mi355x.so`` (gfx950 HIP kernels
plus a C++ host layer, declared in ``include/bmx_mi355x.h``).  This module is
a thin ctypes layer over it for tests, benchmarks and scripting; it never
falls back to a CPU implementation: if the library or a GPU is missing the
calls raise.

Particle records use the reference's AMReX ``Particle<34,35>`` layout
(src/des/bmx_pc.H:86-96): ``RECORD_DTYPE`` below.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional, Sequence

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libbmx_mi355x.so")

# --- record layout (src/chemistry/bmx_chem.H:38-118) -----------------------
NREAL = 34
NINT = 35
RECORD_DTYPE = np.dtype(
    {
        "names": ["pos", "rdata", "id", "cpu", "idata"],
        "formats": [("<f8", 3), ("<f8", NREAL), "<i4", "<i4", ("<i4", NINT)],
        "offsets": [0, 24, 296, 300, 304],
        "itemsize": 448,
    }
)


class R:  # realIdx
    radius, c_length, theta, phi, area, vol = 0, 1, 2, 3, 4, 5
    velx, vely, velz, wx, wy, wz = 6, 7, 8, 9, 10, 11
    fx, fy, fz, taux, tauy, tauz, gx, gy, gz, dadt, dvdt = range(12, 23)
    tau_split, bond_scale, first_data = 23, 24, 25


class I:  # intIdx
    cell_type, n_bnds = 5, 6
    seg1_id1, seg1_id2, site1 = 7, 11, 15
    position, fuse_flag, split_flag, new_flag = 19, 20, 21, 22
    fuse_id, fuse_cpu, fuse_tip = 23, 24, 25
    dltd_site1, dltd_cpu2, fix_site, id, cpu = 26, 31, 32, 33, 34


YEAST, FUNGI = 0, 1
TIP, SECOND_1, SECOND_2, INTERIOR = 0, 1, 2, 3

# fpar order (src/chemistry/bmx_cell_interaction.cpp:84-93)
FPAR_NAMES = (
    "boundary_width", "stiffness", "z_wall", "wall_boundary_width",
    "wall_stiffness", "gravity", "bond_strength", "bond_cutoff", "viscous_drag",
)

NBR_VERLET, NBR_STENCIL = 0, 1
AOS_HOST, AOS_DEVICE, AOS_IN_PLACE = 0, 1, 2
FIELD_POSITIONS, FIELD_SIZES, FIELD_ORIENTATIONS, FIELD_BONDS, FIELD_SPECIES = \
    1, 2, 4, 8, 16
# limits of bmx_evolve_aos_noalloc (include/bmx_mi355x.h)
BMX_NOALLOC_MAX_PARTICLES = 32768
BMX_NOALLOC_MAX_NEIGHBORS = 64
BMX_NOALLOC_MAX_REBUILD_EVERY = 64
DEPOSIT_ONE_TO_ONE, DEPOSIT_TRILINEAR = 0, 1

STATUS = {
    0: "OK", 1: "BAD_ARG", 2: "OOM", 3: "HIP", 4: "NAN", 5: "NBR_OVERFLOW",
    6: "SKIN", 7: "STATE", 8: "IO", 9: "BOND_OVERFLOW", 10: "NO_DEVICE",
    11: "MESH", 12: "CHEM",
}


class BMXError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{STATUS.get(code, code)}: {msg}")
        self.code = code


# --- C structs ---------------------------------------------------------------
class Layout(C.Structure):
    _fields_ = [(n, C.c_int32) for n in (
        "stride_bytes", "pos_offset", "rdata_offset", "id_offset",
        "cpu_offset", "idata_offset", "n_real", "n_int")]


class Geometry(C.Structure):
    _fields_ = [
        ("prob_lo", C.c_double * 3), ("prob_hi", C.c_double * 3),
        ("periodic", C.c_int32 * 3), ("n_cell", C.c_int32 * 3),
    ]


class NeighborParams(C.Structure):
    _fields_ = [
        ("mode", C.c_int32), ("rebuild_every", C.c_int32),
        ("max_neighbors", C.c_int32), ("bin_factor_pct", C.c_int32),
        ("skin", C.c_double), ("neighborhood", C.c_double),
    ]


class LaunchParams(C.Structure):
    _fields_ = [("block_size", C.c_int32), ("xcd_remap", C.c_int32)]


class Timing(C.Structure):
    _fields_ = [
        ("force_ms", C.c_double), ("force_launches", C.c_int64),
        ("rebuild_ms", C.c_double), ("rebuilds", C.c_int64),
        ("neighbor_entries", C.c_int64), ("max_row", C.c_int32),
        ("nbins", C.c_int32), ("rebuilds_total", C.c_int64),
        ("steps_total", C.c_int64), ("guard_redone", C.c_int64),
    ]


class AdjointTiming(C.Structure):
    _fields_ = [("forward_ms", C.c_double), ("backward_ms", C.c_double),
                ("segments", C.c_int32), ("pad", C.c_int32)]


class CheckpointHeader(C.Structure):
    _fields_ = [("version", C.c_int32), ("n_real", C.c_int32), ("n_int", C.c_int32),
                ("finest_level", C.c_int32), ("n_particles", C.c_int64),
                ("next_id", C.c_int64)]


class Tuning(C.Structure):
    _fields_ = [("block_size", C.c_int32), ("xcd_remap", C.c_int32),
                ("bin_factor_pct", C.c_int32), ("from_cache", C.c_int32),
                ("us_per_step", C.c_double), ("candidates", C.c_int32),
                ("pad", C.c_int32)]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "pad"}


_lib: Optional[C.CDLL] = None


def lib_path() -> str:
    return LIB_PATH


def load_library() -> C.CDLL:
    """Load libbmx_mi355x.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is missing; build it with `python -m boltzmannmfx_amd.build`")
    lib = C.CDLL(LIB_PATH)
    P, vp = C.POINTER, C.c_void_p
    sig = {
        "bmx_abi_version": (C.c_int32, []),
        "bmx_last_error": (C.c_char_p, []),
        "bmx_ctx_create": (C.c_int, [P(vp), C.c_int32]),
        "bmx_ctx_destroy": (None, [vp]),
        "bmx_set_force_params": (C.c_int, [vp, P(C.c_double)]),
        "bmx_set_mobility": (C.c_int, [vp, C.c_double]),
        "bmx_set_geometry": (C.c_int, [vp, P(Geometry)]),
        "bmx_set_neighbor_params": (C.c_int, [vp, P(NeighborParams)]),
        "bmx_set_launch_params": (C.c_int, [vp, P(LaunchParams)]),
        "bmx_default_layout": (Layout, []),
        "bmx_bind_particles_aos": (C.c_int, [vp, vp, C.c_int64, P(Layout), C.c_int32]),
        "bmx_bind_colonies_aos": (C.c_int, [vp, vp, P(C.c_int64), C.c_int32,
                                            P(Geometry), P(Layout), C.c_int32]),
        "bmx_sync_to_aos": (C.c_int, [vp, vp]),
        "bmx_num_particles": (C.c_int64, [vp]),
        "bmx_init_bonds": (C.c_int, [vp]),
        "bmx_evolve": (C.c_int, [vp, C.c_double, C.c_int32]),
        "bmx_evolve_graph": (C.c_int, [vp, C.c_double, C.c_int32]),
        "bmx_get_fc": (C.c_int, [vp, P(C.c_double)]),
        "bmx_get_positions": (C.c_int, [vp, P(C.c_double)]),
        "bmx_get_orientations": (C.c_int, [vp, P(C.c_double)]),
        "bmx_get_int_fields": (C.c_int, [vp, P(C.c_int32)]),
        "bmx_get_neighbors": (C.c_int, [vp, P(C.c_int64), P(C.c_int32), C.c_int64,
                                        P(C.c_int64)]),
        "bmx_get_bins": (C.c_int, [vp, P(C.c_int32)]),
        "bmx_get_bin_dims": (C.c_int, [vp, C.c_int32, P(C.c_int32)]),
        "bmx_set_timing": (C.c_int, [vp, C.c_int32]),
        "bmx_get_timing": (C.c_int, [vp, P(Timing)]),
        "bmx_reset_timing": (C.c_int, [vp]),
        "bmx_synchronize": (C.c_int, [vp]),
        "bmx_adjoint": (C.c_int, [vp, C.c_double, C.c_int32, C.c_int32, P(C.c_double),
                                  P(C.c_double), P(C.c_double), P(C.c_double),
                                  P(AdjointTiming)]),
        "bmx_exchange": (C.c_int, [vp, P(C.c_double), C.c_double]),
        "bmx_split_particles": (C.c_int, [vp, P(C.c_double), C.c_uint64, C.c_uint32,
                                          C.c_int32, C.c_int32, P(C.c_int64)]),
        "bmx_fusion": (C.c_int, [vp, P(C.c_double), C.c_uint64, C.c_uint32, C.c_int32,
                                 C.c_int32, P(C.c_int32), P(C.c_int64)]),
        "bmx_calc_volume_fraction": (C.c_int, [vp, C.c_int32, vp, C.c_int32]),
        "bmx_calc_txfr_fluid": (C.c_int, [vp, C.c_int32, vp, C.c_int32]),
        "bmx_count_particles_per_cell": (C.c_int, [vp, vp, C.c_int32]),
        "bmx_calc_txfr_particle": (C.c_int, [vp, P(C.c_double), vp, vp, C.c_double,
                                             C.c_int32, C.c_int32, C.c_int32,
                                             C.c_int32]),
        "bmx_checkpoint_info": (C.c_int, [C.c_char_p, P(CheckpointHeader)]),
        "bmx_read_checkpoint": (C.c_int, [C.c_char_p, vp, C.c_int64, P(C.c_int64),
                                          P(Layout), P(C.c_int32), C.c_int32,
                                          P(C.c_int32), C.c_int32]),
        "bmx_write_checkpoint": (C.c_int, [C.c_char_p, vp, C.c_int64, P(Layout),
                                           P(Geometry), C.c_int32, C.c_int64]),
        "bmx_autotune": (C.c_int, [vp, C.c_double, C.c_int32, C.c_char_p, P(Tuning)]),
        "bmx_bind_tiles_aos": (C.c_int, [vp, P(vp), P(C.c_int64), C.c_int32,
                                         P(Layout), C.c_int32]),
        "bmx_refresh_from_aos": (C.c_int, [vp, vp, C.c_uint32]),
        "bmx_evolve_async": (C.c_int, [vp, C.c_double, C.c_int32]),
        "bmx_wait": (C.c_int, [vp]),
        "bmx_query": (C.c_int, [vp, P(C.c_int32)]),
        "bmx_set_stream": (C.c_int, [vp, vp]),
        "bmx_set_guard": (C.c_int, [vp, C.c_int32]),
        "bmx_evolve_aos_noalloc": (C.c_int, [vp, C.c_int64, P(Layout), P(Geometry),
                                             P(C.c_double), P(NeighborParams),
                                             C.c_double, C.c_int32, vp]),
        "bmx_read_particles_ascii": (C.c_int, [C.c_char_p, vp, C.c_int64,
                                               P(C.c_int64), P(Layout)]),
        "bmx_write_particles_ascii": (C.c_int, [C.c_char_p, vp, C.c_int64, P(Layout)]),
        "bmx_mean_sphere_volume": (C.c_double, [C.c_double, C.c_double]),
        "bmx_generate_spheres": (C.c_int, [vp, C.c_int64, C.c_double, C.c_double,
                                           C.c_double, C.c_uint64, P(C.c_double),
                                           P(C.c_int32), C.c_int32, P(C.c_double),
                                           P(C.c_double), P(Layout)]),
        "bmx_generate_hyphae": (C.c_int, [vp, C.c_int64, C.c_int32, C.c_double,
                                          C.c_double, C.c_double, C.c_double,
                                          C.c_double, C.c_uint64, P(C.c_double),
                                          P(C.c_double), C.c_double, P(Layout)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class _DLDevice(C.Structure):
    _fields_ = [("device_type", C.c_int32), ("device_id", C.c_int32)]


class _DLDataType(C.Structure):
    _fields_ = [("code", C.c_uint8), ("bits", C.c_uint8), ("lanes", C.c_uint16)]


class _DLTensor(C.Structure):
    _fields_ = [("data", C.c_void_p), ("device", _DLDevice), ("ndim", C.c_int32),
                ("dtype", _DLDataType), ("shape", C.POINTER(C.c_int64)),
                ("strides", C.POINTER(C.c_int64)), ("byte_offset", C.c_uint64)]


_KDL_ROCM = 10   # DLDeviceType kDLROCM
_KDL_FLOAT = 2   # DLDataTypeCode kDLFloat


def _dlpack_view(a, device_type: int = _KDL_ROCM):
    """(address, element count) of a contiguous float64 array through the
    DLPack protocol, None for an object without it or on another device
    (ROCm, unless `device_type` says otherwise).
    The capsule is read, not consumed: its destructor hands the tensor back
    to its producer, and the address stays valid while `a` lives."""
    dev = getattr(a, "__dlpack_device__", None)
    if dev is None or not hasattr(a, "__dlpack__"):
        return None
    if int(dev()[0]) != device_type:
        return None
    cap = a.__dlpack__()
    get = C.pythonapi.PyCapsule_GetPointer
    get.restype = C.c_void_p
    get.argtypes = [C.py_object, C.c_char_p]
    t = _DLTensor.from_address(get(cap, b"dltensor"))  # DLManagedTensor starts with it
    assert t.dtype.code == _KDL_FLOAT and t.dtype.bits == 64 and t.dtype.lanes == 1, \
        "device meshes are float64"
    shape = [t.shape[i] for i in range(t.ndim)]
    if t.strides:
        want = 1
        for i in reversed(range(t.ndim)):
            assert shape[i] == 1 or t.strides[i] == want, "device meshes must be contiguous"
            want *= shape[i]
    return int(t.data) + int(t.byte_offset), int(np.prod(shape, dtype=np.int64))


def _device_view(a):
    """(address, element count) of a contiguous float64 device array, taken
    through DLPack (`__dlpack__` on the ROCm device) or
    `__cuda_array_interface__`; None for a host array."""
    dl = _dlpack_view(a)
    if dl is not None:
        return dl
    cai = getattr(a, "__cuda_array_interface__", None)
    if cai is None:
        return None
    assert cai["typestr"] == "<f8", "device meshes are float64"
    assert cai.get("strides") is None, "device meshes must be contiguous"
    return int(cai["data"][0]), int(np.prod(cai["shape"]))


def _check(code: int) -> None:
    if code != 0:
        msg = load_library().bmx_last_error().decode(errors="replace")
        raise BMXError(code, msg)


def evolve_aos_noalloc(aos_device: int, n: int, geometry: Geometry, fpar,
                       dt: float, nsteps: int, mode=NBR_VERLET, rebuild_every=10,
                       skin=0.0, max_neighbors=0, bin_factor_pct=100,
                       neighborhood=1.0, stream: int = 0) -> None:
    """bmx_evolve_aos_noalloc: bind + evolve + sync on n device records at
    address aos_device, in place, with no device allocation."""
    lib = load_library()
    fp = np.ascontiguousarray(fpar, dtype=np.float64)
    assert fp.shape == (9,)
    p = NeighborParams(int(mode), int(rebuild_every), int(max_neighbors),
                       int(bin_factor_pct), float(skin), float(neighborhood))
    lay = lib.bmx_default_layout()
    _check(lib.bmx_evolve_aos_noalloc(C.c_void_p(aos_device), int(n), C.byref(lay),
                                      C.byref(geometry), _dptr(fp), C.byref(p),
                                      float(dt), int(nsteps),
                                      C.c_void_p(stream) if stream else None))


def make_geometry(prob_lo, prob_hi, periodic, n_cell=(1, 1, 1)) -> Geometry:
    g = Geometry()
    for d in range(3):
        g.prob_lo[d] = float(prob_lo[d])
        g.prob_hi[d] = float(prob_hi[d])
        g.periodic[d] = int(bool(periodic[d]))
        g.n_cell[d] = int(n_cell[d])
    return g


class Context:
    """One bmx_ctx: device buffers, stream and the bound particles."""

    def __init__(self, device: int = 0):
        self.lib = load_library()
        h = C.c_void_p()
        _check(self.lib.bmx_ctx_create(C.byref(h), int(device)))
        self.h = h
        self.n = 0

    def close(self) -> None:
        if self.h:
            self.lib.bmx_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # configuration
    def set_force_params(self, fpar: Sequence[float]) -> None:
        a = np.ascontiguousarray(fpar, dtype=np.float64)
        assert a.shape == (9,)
        _check(self.lib.bmx_set_force_params(self.h, _dptr(a)))

    def set_mobility(self, m: float) -> None:
        _check(self.lib.bmx_set_mobility(self.h, float(m)))

    def set_geometry(self, prob_lo, prob_hi, periodic, n_cell=(1, 1, 1)) -> None:
        g = make_geometry(prob_lo, prob_hi, periodic, n_cell)
        _check(self.lib.bmx_set_geometry(self.h, C.byref(g)))
        self._n_cell = tuple(int(v) for v in n_cell)

    # particle <-> mesh.  Host meshes are numpy arrays [comp, k, j, i];
    # device meshes are any contiguous float64 array on the context's device
    # that speaks DLPack or __cuda_array_interface__ (a torch tensor, for
    # one), passed by address.
    def _mesh_arg(self, a, count: int):
        dv = _device_view(a)
        if dv is not None:
            assert dv[1] == count, "device mesh has the wrong size"
            return C.c_void_p(dv[0]), 1
        assert a.flags.c_contiguous and a.dtype == np.float64 and a.size == count
        return a.ctypes.data_as(C.c_void_p), 0

    def _cells(self, ghost: int = 0) -> int:
        nx, ny, nz = self._n_cell
        return (nx + 2 * ghost) * (ny + 2 * ghost) * (nz + 2 * ghost)

    def _mesh_out(self, ncomp: int, out):
        nx, ny, nz = self._n_cell
        if out is None:
            out = np.zeros((ncomp, nz, ny, nx)) if ncomp > 1 else np.zeros((nz, ny, nx))
        return out

    def volume_fraction(self, scheme: int = DEPOSIT_ONE_TO_ONE, out=None):
        """bmx_calc_volume_fraction: vf per cell, shape (nz, ny, nx)."""
        out = self._mesh_out(1, out)
        p, dev = self._mesh_arg(out, self._cells())
        _check(self.lib.bmx_calc_volume_fraction(self.h, int(scheme), p, dev))
        return out

    def txfr_fluid(self, scheme: int = DEPOSIT_ONE_TO_ONE, out=None):
        """bmx_calc_txfr_fluid: rdata[31..33] summed per cell, (3, nz, ny, nx)."""
        out = self._mesh_out(3, out)
        p, dev = self._mesh_arg(out, 3 * self._cells())
        _check(self.lib.bmx_calc_txfr_fluid(self.h, int(scheme), p, dev))
        return out

    def count_per_cell(self, out=None):
        out = self._mesh_out(1, out)
        p, dev = self._mesh_arg(out, self._cells())
        _check(self.lib.bmx_count_particles_per_cell(self.h, p, dev))
        return out

    def split_particles(self, params, seed: int, call: int, next_id: int,
                        cpu: int = 0) -> np.ndarray:
        """split_particles: divides, grows new tips and branches.  params =
        (max_vol, max_seg_length, max_seg_radius, branching_probability,
        splitting_probability, overlap, seg_split_length).  Returns the
        (possibly grown) record array, now the context's bound buffer."""
        sp = np.ascontiguousarray(params, dtype=np.float64)
        assert sp.size == 7
        nn = C.c_int64()
        _check(self.lib.bmx_split_particles(self.h, _dptr(sp), int(seed), int(call),
                                            int(next_id), int(cpu), C.byref(nn)))
        if nn.value == 0:
            return self._recs
        m = int(self.lib.bmx_num_particles(self.h))
        out = np.zeros(m, dtype=RECORD_DTYPE)
        _check(self.lib.bmx_sync_to_aos(self.h, out.ctypes.data_as(C.c_void_p)))
        self._recs = out
        self.n = m
        return out

    def fusion(self, params, seed: int, call: int, next_id: int, cpu: int = 0):
        """Tip fusion, interior fusion and cleanup.  params =
        (fusion_probability, max_fusion_separation, bond_scaling_increment,
        max_seg_length).  Returns (records, fused, new segments); the
        records are the context's bound buffer afterwards."""
        fp = np.ascontiguousarray(params, dtype=np.float64)
        assert fp.size == 4
        fused, nn = C.c_int32(), C.c_int64()
        _check(self.lib.bmx_fusion(self.h, _dptr(fp), int(seed), int(call), int(next_id),
                                   int(cpu), C.byref(fused), C.byref(nn)))
        if nn.value > 0:
            m = int(self.lib.bmx_num_particles(self.h))
            self._recs = np.zeros(m, dtype=RECORD_DTYPE)
            self.n = m
        self.sync()
        return self._recs, bool(fused.value), int(nn.value)

    def txfr_particle(self, chempar, x_k, vf, dt: float, nloop: int = 4,
                      cnc_scheme: int = DEPOSIT_ONE_TO_ONE,
                      vf_scheme: int = DEPOSIT_ONE_TO_ONE) -> None:
        """bmx_calc_txfr_particle: x_k (3, nz+2, ny+2, nx+2) and vf
        (nz+2, ny+2, nx+2) with one ghost layer; updates the records (read
        back by sync()).  Both host or both device arrays."""
        cp = np.ascontiguousarray(chempar, dtype=np.float64)
        assert cp.size == 10
        if _device_view(x_k) is None:
            x_k = np.ascontiguousarray(x_k, dtype=np.float64)
            vf = np.ascontiguousarray(vf, dtype=np.float64)
        px, dx_ = self._mesh_arg(x_k, 3 * self._cells(1))
        pv, dv_ = self._mesh_arg(vf, self._cells(1))
        assert dx_ == dv_, "x_k and vf must both be host or both device arrays"
        _check(self.lib.bmx_calc_txfr_particle(
            self.h, _dptr(cp), px, pv, float(dt), int(nloop), int(cnc_scheme),
            int(vf_scheme), dx_))

    def set_neighbor_params(self, mode=NBR_VERLET, rebuild_every=10, skin=0.0,
                            max_neighbors=0, bin_factor_pct=100,
                            neighborhood=1.0) -> None:
        p = NeighborParams(int(mode), int(rebuild_every), int(max_neighbors),
                           int(bin_factor_pct), float(skin), float(neighborhood))
        _check(self.lib.bmx_set_neighbor_params(self.h, C.byref(p)))

    def set_launch_params(self, block_size=64, xcd_remap=1) -> None:
        p = LaunchParams(int(block_size), int(xcd_remap))
        _check(self.lib.bmx_set_launch_params(self.h, C.byref(p)))

    # particles
    def bind(self, recs: np.ndarray) -> None:
        assert recs.dtype == RECORD_DTYPE and recs.flags.c_contiguous
        self._recs = recs
        lay = self.lib.bmx_default_layout()
        _check(self.lib.bmx_bind_particles_aos(
            self.h, recs.ctypes.data_as(C.c_void_p), len(recs), C.byref(lay), 0))
        self.n = len(recs)

    def bind_device(self, address: int, n: int, in_place: bool = False) -> None:
        """Binds n records in device memory at `address` (copied, or used in
        place)."""
        lay = self.lib.bmx_default_layout()
        self._recs = None
        _check(self.lib.bmx_bind_particles_aos(
            self.h, C.c_void_p(address), int(n), C.byref(lay),
            AOS_IN_PLACE if in_place else AOS_DEVICE))
        self.n = int(n)

    def bind_tiles(self, tiles: Sequence[np.ndarray]) -> None:
        """Several host record arrays as one neighbour domain (the context's
        geometry); sync() writes each back."""
        for t in tiles:
            assert t.dtype == RECORD_DTYPE and t.flags.c_contiguous
        self._tiles = list(tiles)
        self._recs = None
        ptrs = (C.c_void_p * len(tiles))(*[t.ctypes.data for t in tiles])
        counts = np.array([len(t) for t in tiles], dtype=np.int64)
        lay = self.lib.bmx_default_layout()
        _check(self.lib.bmx_bind_tiles_aos(
            self.h, ptrs, counts.ctypes.data_as(C.POINTER(C.c_int64)), len(tiles),
            C.byref(lay), 0))
        self.n = int(counts.sum())

    def refresh(self, fields: int, recs: Optional[np.ndarray] = None) -> None:
        """bmx_refresh_from_aos: re-read the given BMX_FIELD_* bits."""
        ptr = recs.ctypes.data_as(C.c_void_p) if recs is not None else None
        _check(self.lib.bmx_refresh_from_aos(self.h, ptr, int(fields)))

    def evolve_async(self, dt: float, nsteps: int) -> None:
        _check(self.lib.bmx_evolve_async(self.h, float(dt), int(nsteps)))

    def wait(self) -> None:
        _check(self.lib.bmx_wait(self.h))

    def query(self) -> bool:
        d = C.c_int32()
        _check(self.lib.bmx_query(self.h, C.byref(d)))
        return bool(d.value)

    def set_guard(self, on: bool = True) -> None:
        """Guarded Verlet steps: periods that exceed the skin are taken again
        with per-step lists instead of failing."""
        _check(self.lib.bmx_set_guard(self.h, 1 if on else 0))

    def set_stream(self, stream: int = 0) -> None:
        _check(self.lib.bmx_set_stream(self.h, C.c_void_p(stream) if stream else None))

    def bind_colonies(self, recs: np.ndarray, colony_start: Sequence[int],
                      geoms: Sequence[Geometry]) -> None:
        assert recs.dtype == RECORD_DTYPE and recs.flags.c_contiguous
        self._recs = recs
        cs = np.ascontiguousarray(colony_start, dtype=np.int64)
        ga = (Geometry * len(geoms))(*geoms)
        lay = self.lib.bmx_default_layout()
        _check(self.lib.bmx_bind_colonies_aos(
            self.h, recs.ctypes.data_as(C.c_void_p),
            cs.ctypes.data_as(C.POINTER(C.c_int64)), len(geoms), ga,
            C.byref(lay), 0))
        self.n = len(recs)

    def sync(self, recs: Optional[np.ndarray] = None) -> None:
        if recs is None:
            recs = getattr(self, "_recs", None)
        ptr = recs.ctypes.data_as(C.c_void_p) if recs is not None else None
        _check(self.lib.bmx_sync_to_aos(self.h, ptr))

    # hot path
    def init_bonds(self) -> None:
        _check(self.lib.bmx_init_bonds(self.h))

    def evolve(self, dt: float, nsteps: int) -> None:
        _check(self.lib.bmx_evolve(self.h, float(dt), int(nsteps)))

    def evolve_graph(self, dt: float, nsteps: int) -> None:
        _check(self.lib.bmx_evolve_graph(self.h, float(dt), int(nsteps)))

    def adjoint(self, dt: float, nsteps: int, target: np.ndarray,
                checkpoint_every: int = 20):
        """(loss, [dL/dmu, dL/dW_A, dL/dgamma], dL/dradius, timing) of nsteps
        sphere steps from the current state; the context ends in the final
        state.  target: (n, 3) positions in original order."""
        tg = np.ascontiguousarray(target, dtype=np.float64).reshape(-1)
        assert tg.size == 3 * self.n
        loss = C.c_double()
        g3 = np.zeros(3)
        gr = np.zeros(max(self.n, 1))
        tm = AdjointTiming()
        _check(self.lib.bmx_adjoint(self.h, float(dt), int(nsteps), int(checkpoint_every),
                                    _dptr(tg), C.byref(loss), _dptr(g3), _dptr(gr),
                                    C.byref(tm)))
        return loss.value, g3, gr[: self.n], tm

    def exchange(self, xpar, dt: float) -> None:
        """ParticleExchange between bonded segments (species in rdata
        25..27, increments in 28..30; written back by sync())."""
        xp = np.ascontiguousarray(xpar, dtype=np.float64)
        assert xp.size == 3
        _check(self.lib.bmx_exchange(self.h, _dptr(xp), float(dt)))

    def autotune(self, dt: float, trial_steps: int = 20,
                 cache: Optional[str] = None) -> dict:
        """Times the launch shapes and bin edges on the bound particles,
        applies the fastest and returns it; the state is left unchanged.
        Trials of several rebuild periods, after the first rebuild, rank the
        shapes as a long run does."""
        t = Tuning()
        _check(self.lib.bmx_autotune(self.h, float(dt), int(trial_steps),
                                     cache.encode() if cache else None, C.byref(t)))
        return t.as_dict()

    # inspection
    def fc(self) -> np.ndarray:
        out = np.empty(6 * self.n, dtype=np.float64)
        _check(self.lib.bmx_get_fc(self.h, _dptr(out)))
        return out.reshape(6, self.n)

    def positions(self) -> np.ndarray:
        out = np.empty(3 * self.n, dtype=np.float64)
        _check(self.lib.bmx_get_positions(self.h, _dptr(out)))
        return out.reshape(self.n, 3)

    def orientations(self) -> np.ndarray:
        out = np.empty(2 * self.n, dtype=np.float64)
        _check(self.lib.bmx_get_orientations(self.h, _dptr(out)))
        return out.reshape(self.n, 2)

    def int_fields(self) -> np.ndarray:
        out = np.empty(35 * self.n, dtype=np.int32)
        _check(self.lib.bmx_get_int_fields(
            self.h, out.ctypes.data_as(C.POINTER(C.c_int32))))
        return out.reshape(self.n, 35)

    def neighbors(self):
        off = np.empty(self.n + 1, dtype=np.int64)
        tot = C.c_int64()
        _check(self.lib.bmx_get_neighbors(
            self.h, off.ctypes.data_as(C.POINTER(C.c_int64)), None, 0, C.byref(tot)))
        idx = np.empty(max(tot.value, 1), dtype=np.int32)
        _check(self.lib.bmx_get_neighbors(
            self.h, off.ctypes.data_as(C.POINTER(C.c_int64)),
            idx.ctypes.data_as(C.POINTER(C.c_int32)), len(idx), C.byref(tot)))
        return off, idx[: tot.value]

    def bins(self) -> np.ndarray:
        out = np.empty(self.n, dtype=np.int32)
        _check(self.lib.bmx_get_bins(self.h, out.ctypes.data_as(C.POINTER(C.c_int32))))
        return out

    def bin_dims(self, colony: int = 0):
        out = (C.c_int32 * 3)()
        _check(self.lib.bmx_get_bin_dims(self.h, int(colony), out))
        return tuple(int(v) for v in out)

    def set_timing(self, on) -> None:
        """Events around every rebuild and every run of force launches
        between rebuilds."""
        _check(self.lib.bmx_set_timing(self.h, 1 if on else 0))

    def timing(self) -> Timing:
        t = Timing()
        _check(self.lib.bmx_get_timing(self.h, C.byref(t)))
        return t

    def reset_timing(self) -> None:
        _check(self.lib.bmx_reset_timing(self.h))

    def synchronize(self) -> None:
        _check(self.lib.bmx_synchronize(self.h))


# --- particle files and synthetic inputs ------------------------------------
def read_particles_ascii(path: str) -> np.ndarray:
    """InitParticlesAscii's 27-field format -> records (src/des/bmx_pc_init.cpp:17-132)."""
    lib = load_library()
    n = C.c_int64()
    _check(lib.bmx_read_particles_ascii(path.encode(), None, 0, C.byref(n), None))
    recs = np.zeros(n.value, dtype=RECORD_DTYPE)
    _check(lib.bmx_read_particles_ascii(path.encode(), recs.ctypes.data_as(C.c_void_p),
                                        n.value, C.byref(n), None))
    return recs


def write_particles_ascii(path: str, recs: np.ndarray) -> None:
    lib = load_library()
    recs = np.ascontiguousarray(recs, dtype=RECORD_DTYPE)
    _check(lib.bmx_write_particles_ascii(path.encode(), recs.ctypes.data_as(C.c_void_p),
                                         len(recs), None))


def generate_spheres(n: int, rmin: float, rmax: float, phi: float, seed: int,
                     aspect=(2.0, 2.0, 1.0), periodic=(1, 1, 0),
                     overlap_free: bool = True):
    """Seeded sphere colony; returns (records, prob_lo, prob_hi)."""
    lib = load_library()
    recs = np.zeros(n, dtype=RECORD_DTYPE)
    asp = np.asarray(aspect, dtype=np.float64)
    per = np.asarray(periodic, dtype=np.int32)
    lo = np.zeros(3)
    hi = np.zeros(3)
    _check(lib.bmx_generate_spheres(
        recs.ctypes.data_as(C.c_void_p), n, rmin, rmax, phi, seed & (2**64 - 1),
        _dptr(asp), per.ctypes.data_as(C.POINTER(C.c_int32)), int(overlap_free),
        _dptr(lo), _dptr(hi), None))
    return recs, lo, hi


def generate_hyphae(n: int, nroots: int, seg_len: float, seg_rad: float,
                    branch_prob: float, angle_min: float, angle_max: float,
                    seed: int, box_lo, box_hi, z_surface: float) -> np.ndarray:
    """Seeded branching segment chains lying above z_surface."""
    lib = load_library()
    recs = np.zeros(n, dtype=RECORD_DTYPE)
    lo = np.asarray(box_lo, dtype=np.float64)
    hi = np.asarray(box_hi, dtype=np.float64)
    _check(lib.bmx_generate_hyphae(
        recs.ctypes.data_as(C.c_void_p), n, nroots, seg_len, seg_rad, branch_prob,
        angle_min, angle_max, seed & (2**64 - 1), _dptr(lo), _dptr(hi),
        float(z_surface), None))
    return recs


def mean_sphere_volume(rmin: float, rmax: float) -> float:
    return load_library().bmx_mean_sphere_volume(rmin, rmax)


# --- particle checkpoints -----------------------------------------------------
def checkpoint_info(path: str) -> dict:
    """Header of a particle checkpoint directory (the "particles" part of a
    BMX checkpoint)."""
    h = CheckpointHeader()
    _check(load_library().bmx_checkpoint_info(path.encode(), C.byref(h)))
    return {k: getattr(h, k) for k, _ in h._fields_}


def _legacy_maps():
    """Component maps of the two older record layouts in the checkpoints the
    reference ships, onto the current Particle<34,35> record.  Inferred from
    the decoded files (field values and the record layout of
    src/chemistry/bmx_chem.H:38-118), not from the older sources, which the
    reference does not hold.

    31 reals / 5 ints (exec/rst_test/chk20000, exec/nlev_large_test/dbg/
    chk300000, spheres): reals = ellipsoid axes a, b, c, Euler angles psi,
    theta, phi, area, vol, vel[3], w[3], f[3], tau[3], two rates, then the
    9 species values; ints = the chemistry header only (all cells yeast).

    33 reals / 33 ints (exec/fungi/chk357900, segments): the current reals
    without bond_scale (species start at 24), the current ints without
    fuse_tip and fix_site."""
    r31 = [0, 0, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 17, 18, 19,
           -1, -1, -1, 20, 21, -1, -1] + list(range(22, 31))
    i5 = [0, 1, 2, 3, 4]
    r33 = list(range(0, 24)) + [-1] + list(range(24, 33))
    i33 = list(range(0, 25)) + [-1] + list(range(25, 31)) + [-1, 31, 32]
    return {(31, 5): (r31, i5), (33, 33): (r33, i33)}


LEGACY_CHECKPOINT_MAPS = _legacy_maps()


def read_checkpoint(path: str, real_map=None, int_map=None) -> np.ndarray:
    """Particles of a checkpoint directory as Particle<34,35> records.  The
    current layout (34 reals, 35 ints) is read as is; the older layouts of
    the shipped checkpoints go through LEGACY_CHECKPOINT_MAPS, after which
    bond_scale is 1 where the file had none and idata id/cpu are the
    particle's id words, as InitParticlesAscii sets them."""
    lib = load_library()
    info = checkpoint_info(path)
    legacy = False
    if real_map is None and int_map is None and (info["n_real"], info["n_int"]) != (34, 35):
        key = (info["n_real"], info["n_int"])
        if key not in LEGACY_CHECKPOINT_MAPS:
            raise BMXError(1, f"no component map for {key[0]} reals / {key[1]} ints")
        real_map, int_map = LEGACY_CHECKPOINT_MAPS[key]
        legacy = True
    recs = np.zeros(max(info["n_particles"], 1), dtype=RECORD_DTYPE)
    n = C.c_int64()
    lay = lib.bmx_default_layout()

    def arr(m):
        if m is None:
            return None, 0
        a = np.ascontiguousarray(m, dtype=np.int32)
        return a, a.size
    rm, nr = arr(real_map)
    im, ni = arr(int_map)
    _check(lib.bmx_read_checkpoint(
        path.encode(), recs.ctypes.data_as(C.c_void_p), len(recs), C.byref(n),
        C.byref(lay), rm.ctypes.data_as(C.POINTER(C.c_int32)) if rm is not None else None,
        nr, im.ctypes.data_as(C.POINTER(C.c_int32)) if im is not None else None, ni))
    recs = recs[: n.value].copy()
    if legacy:
        rm_l = list(real_map)
        if rm_l[R.bond_scale] < 0:
            recs["rdata"][:, R.bond_scale] = 1.0
        recs["idata"][:, I.id] = recs["id"]
        recs["idata"][:, I.cpu] = recs["cpu"]
    return recs


def write_checkpoint(path: str, recs: np.ndarray, geometry: "Geometry", level: int = 0,
                     next_id: Optional[int] = None) -> None:
    """Writes recs as a Version_Two_Dot_One_double particle checkpoint with
    one grid spanning geometry.n_cell at `level`."""
    assert recs.dtype == RECORD_DTYPE and recs.flags.c_contiguous
    lib = load_library()
    lay = lib.bmx_default_layout()
    nid = int(next_id) if next_id is not None else (int(recs["id"].max()) + 1 if len(recs) else 1)
    _check(lib.bmx_write_checkpoint(path.encode(), recs.ctypes.data_as(C.c_void_p),
                                    len(recs), C.byref(lay), C.byref(geometry),
                                    int(level), nid))
