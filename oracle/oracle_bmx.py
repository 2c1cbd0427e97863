"""ctypes wrapper of the CPU oracle (oracle/bmx_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.  The product library never
loads it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_build", "liboracle_bmx.so")

NREAL, NINT = 34, 35
RECORD_DTYPE = np.dtype(
    {
        "names": ["pos", "rdata", "id", "cpu", "idata"],
        "formats": [("<f8", 3), ("<f8", NREAL), "<i4", "<i4", ("<i4", NINT)],
        "offsets": [0, 24, 296, 300, 304],
        "itemsize": 448,
    }
)
NBR_RADIUS, NBR_STENCIL, NBR_PAIR = 0, 1, 2


class Geom(C.Structure):
    _fields_ = [
        ("plo", C.c_double * 3), ("phi", C.c_double * 3),
        ("periodic", C.c_int32 * 3), ("nbin", C.c_int32 * 3),
    ]


def make_geom(plo, phi, periodic, nbin) -> Geom:
    g = Geom()
    for d in range(3):
        g.plo[d] = float(plo[d])
        g.phi[d] = float(phi[d])
        g.periodic[d] = int(bool(periodic[d]))
        g.nbin[d] = int(nbin[d])
    return g


def build() -> str:
    """Compile the oracle with its Makefile (gcc -O2 -ffp-contract=off)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB


_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        D, I64, I32, vp = C.POINTER(C.c_double), C.c_int64, C.c_int32, C.c_void_p
        P = C.POINTER
        sigs = {
            "orc_max_interaction_distance": (C.c_double, [D, D, P(I32), P(I32), C.c_double]),
            "orc_min_distance": (None, [C.c_double] * 12 + [D, D]),
            "orc_evaluate_force": (None, [D, D, D, P(I32), P(I32), D, D, D, D, D]),
            "orc_surface_force": (None, [D, D, P(I32), D, D, D]),
            "orc_fix_bonds": (C.c_int, [D, D, P(I32), D, P(I32), C.c_double]),
            "orc_update_particle": (None, [vp, D, C.c_double]),
            "orc_wrap": (None, [vp, I64, P(Geom)]),
            "orc_bin_index": (I64, [D, P(Geom), P(I32)]),
            "orc_pair_separation": (None, [D, D, P(Geom), D]),
            "orc_build_neighbors": (C.c_int, [vp, I64, P(Geom), C.c_int, C.c_double,
                                              P(I64), P(P(I32))]),
            "orc_build_neighbors_brute": (C.c_int, [vp, I64, P(Geom), C.c_int,
                                                    C.c_double, P(I64), P(P(I32))]),
            "orc_state_create": (vp, [vp, I64, P(Geom), D, C.c_int, C.c_double, C.c_int]),
            "orc_state_destroy": (None, [vp]),
            "orc_state_evolve": (C.c_int, [vp, C.c_double, C.c_int, D]),
            "orc_state_init_bonds": (C.c_int, [vp]),
            "orc_state_exchange": (C.c_int, [vp, P(C.c_double), C.c_double]),
            "orc_exchange_pair": (None, [P(C.c_double), P(C.c_double), P(I32), P(I32),
                                         P(C.c_double), C.c_double]),
            "orc_state_neighbor_count": (I64, [vp]),
            "orc_state_get_neighbors": (None, [vp, P(I64), P(I32)]),
            "orc_state_step": (I64, [vp]),
            "orc_num_threads": (C.c_int, []),
            "orc_set_pair_wa": (None, [C.c_double]),
            "orc_math": (C.c_double, [C.c_int, C.c_double]),
            "orc_state_set_mobility": (None, [vp, C.c_double]),
            "orc_adjoint": (C.c_int, [vp, I64, P(Geom), D, C.c_double, C.c_double, C.c_int,
                                      C.c_int, C.c_double, D, D, D, D]),
            "orc_set_num_threads": (None, [C.c_int]),
            "orc_deposit": (I64, [vp, I64, P(Geom), C.c_int, C.c_int, C.c_int, D]),
            "orc_volume_fraction": (I64, [vp, I64, P(Geom), C.c_int, D]),
            "orc_count_per_cell": (I64, [vp, I64, P(Geom), D]),
            "orc_chem_update": (C.c_int, [C.c_double, C.c_double, D, D, D, C.c_double,
                                          C.c_int, D, P(I32)]),
            "orc_mesh_to_particle": (C.c_int, [vp, I64, P(Geom), D, D, D, C.c_double,
                                               C.c_int, C.c_int, C.c_int]),
            "orc_pow": (C.c_double, [C.c_double, C.c_double]),
            "orc_pow_dd": (C.c_double, [C.c_double, C.c_double]),
            "orc_pow_third": (C.c_double, [C.c_double, C.c_int]),
            "orc_interp": (None, [D, D, P(Geom), C.c_int, C.c_int, D]),
            "orc_split_particles": (I64, [vp, I64, I64, P(Geom), D, C.c_uint64,
                                          C.c_uint32, C.c_int32, C.c_int32]),
            "orc_fusion": (I64, [vp, I64, I64, P(Geom), D, C.c_int, C.c_double, D,
                                 C.c_uint64, C.c_uint32, C.c_int32, C.c_int32,
                                 P(C.c_int)]),
        }
        for name, (res, args) in sigs.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


libc = C.CDLL(None)
libc.free.argtypes = [C.c_void_p]


def _d(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _i(a):
    return a.ctypes.data_as(C.POINTER(C.c_int32))


def _i64(a):
    return a.ctypes.data_as(C.POINTER(C.c_int64))


# ---- pair / particle kernels --------------------------------------------
def max_interaction_distance(r1, r2, i1, i2, bw) -> float:
    r1 = np.ascontiguousarray(r1, np.float64)
    r2 = np.ascontiguousarray(r2, np.float64)
    i1 = np.ascontiguousarray(i1, np.int32)
    i2 = np.ascontiguousarray(i2, np.int32)
    return lib().orc_max_interaction_distance(_d(r1), _d(r2), _i(i1), _i(i2), bw)


def min_distance(seg_i, seg_j):
    """seg_* = (x1, y1, z1, x2, y2, z2) -> (tau_i, tau_j)"""
    ti, tj = C.c_double(), C.c_double()
    lib().orc_min_distance(*[float(v) for v in seg_i], *[float(v) for v in seg_j],
                           C.byref(ti), C.byref(tj))
    return ti.value, tj.value


def evaluate_force(r12, par1, par2, ipar1, ipar2, fpar):
    """Returns (vel1, vel2, rot1, rot2, ipar1_after, ipar2_after)."""
    r12 = np.ascontiguousarray(r12, np.float64)
    p1 = np.array(par1, np.float64)
    p2 = np.array(par2, np.float64)
    i1 = np.array(ipar1, np.int32)
    i2 = np.array(ipar2, np.int32)
    fp = np.ascontiguousarray(fpar, np.float64)
    out = [np.zeros(3) for _ in range(4)]
    lib().orc_evaluate_force(_d(r12), _d(p1), _d(p2), _i(i1), _i(i2),
                             *[_d(o) for o in out], _d(fp))
    return out[0], out[1], out[2], out[3], i1, i2


def surface_force(pos, par, ipar, fpar):
    pos = np.ascontiguousarray(pos, np.float64)
    p = np.array(par, np.float64)
    ip = np.array(ipar, np.int32)
    fp = np.ascontiguousarray(fpar, np.float64)
    v = np.zeros(3)
    w = np.zeros(3)
    lib().orc_surface_force(_d(pos), _d(p), _i(ip), _d(v), _d(w), _d(fp))
    return v, w, ip


def fix_bonds(r12, rpar_i, ipar_i, rpar_j, ipar_j, rlim=1e-7):
    r12 = np.ascontiguousarray(r12, np.float64)
    ri = np.array(rpar_i, np.float64)
    rj = np.array(rpar_j, np.float64)
    ii = np.array(ipar_i, np.int32)
    ij = np.array(ipar_j, np.int32)
    rc = lib().orc_fix_bonds(_d(r12), _d(ri), _i(ii), _d(rj), _i(ij), rlim)
    return rc, ii, ij


def update_particle(rec: np.ndarray, fc6, dt: float) -> np.ndarray:
    r = np.array(rec, dtype=RECORD_DTYPE).reshape(1).copy()
    f = np.ascontiguousarray(fc6, np.float64)
    lib().orc_update_particle(r.ctypes.data_as(C.c_void_p), _d(f), dt)
    return r[0]


def bin_index(x, geom: Geom):
    x = np.ascontiguousarray(x, np.float64)
    b3 = np.zeros(3, np.int32)
    b = lib().orc_bin_index(_d(x), C.byref(geom), _i(b3))
    return b, b3


def pair_separation(xi, xj, geom: Geom):
    xi = np.ascontiguousarray(xi, np.float64)
    xj = np.ascontiguousarray(xj, np.float64)
    d = np.zeros(3)
    lib().orc_pair_separation(_d(xi), _d(xj), C.byref(geom), _d(d))
    return d


def build_neighbors(recs, geom: Geom, mode: int, radius: float, brute=False, W_A=0.0):
    """CSR (offsets, indices) over particle indices.  For NBR_PAIR, radius is
    the skin and W_A the adhesion band."""
    recs = np.ascontiguousarray(recs, dtype=RECORD_DTYPE)
    lib().orc_set_pair_wa(W_A)
    n = len(recs)
    off = np.zeros(n + 1, np.int64)
    ptr = C.POINTER(C.c_int32)()
    fn = lib().orc_build_neighbors_brute if brute else lib().orc_build_neighbors
    rc = fn(recs.ctypes.data_as(C.c_void_p), n, C.byref(geom), mode, radius,
            _i64(off), C.byref(ptr))
    if rc != 0:
        raise RuntimeError(f"orc_build_neighbors failed: {rc}")
    tot = int(off[n])
    idx = np.ctypeslib.as_array(ptr, shape=(max(tot, 1),))[:tot].copy() if tot else \
        np.zeros(0, np.int32)
    libc.free(C.cast(ptr, C.c_void_p))
    return off, idx


class State:
    """orc_state: the reference's EvolveParticles loop over a record array."""

    def __init__(self, recs: np.ndarray, geom: Geom, fpar, mode: int, radius: float,
                 rebuild_every: int):
        assert recs.dtype == RECORD_DTYPE and recs.flags.c_contiguous
        self.recs = recs
        self.n = len(recs)
        fp = np.ascontiguousarray(fpar, np.float64)
        self.h = lib().orc_state_create(recs.ctypes.data_as(C.c_void_p), self.n,
                                        C.byref(geom), _d(fp), mode, radius,
                                        rebuild_every)
        if not self.h:
            raise MemoryError("orc_state_create failed")

    def close(self):
        if self.h:
            lib().orc_state_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def evolve(self, dt: float, nsteps: int) -> np.ndarray:
        fc = np.zeros(6 * self.n)
        rc = lib().orc_state_evolve(self.h, dt, nsteps, _d(fc))
        if rc != 0:
            raise RuntimeError(f"orc_state_evolve failed: {rc}")
        return fc.reshape(6, self.n)

    def set_mobility(self, m: float) -> None:
        lib().orc_state_set_mobility(self.h, m)

    def init_bonds(self) -> int:
        return lib().orc_state_init_bonds(self.h)

    def exchange(self, xpar, dt: float) -> None:
        """ParticleExchange: rebuild at the current positions, exchange
        between bonded segments, apply the increments (rdata 25..30)."""
        xp = np.ascontiguousarray(xpar, np.float64)
        rc = lib().orc_state_exchange(self.h, _d(xp), float(dt))
        if rc != 0:
            raise RuntimeError(f"orc_state_exchange failed: {rc}")

    def neighbors(self):
        tot = lib().orc_state_neighbor_count(self.h)
        off = np.zeros(self.n + 1, np.int64)
        idx = np.zeros(max(tot, 1), np.int32)
        lib().orc_state_get_neighbors(self.h, _i64(off), _i(idx))
        return off, idx[:tot]


def exchange_pair(rd1, rd2, id1, id2, xpar, dt):
    """evaluateExchange for one pair; rd1/rd2 (34 doubles) are updated."""
    r1 = np.array(rd1, np.float64)  # copies: the caller's arrays stay as given
    r2 = np.array(rd2, np.float64)
    i1 = np.ascontiguousarray(id1, np.int32)
    i2 = np.ascontiguousarray(id2, np.int32)
    xp = np.ascontiguousarray(xpar, np.float64)
    lib().orc_exchange_pair(_d(r1), _d(r2), _i(i1), _i(i2), _d(xp), float(dt))
    return r1, r2


def adjoint(recs, geom: Geom, fpar, mobility, dt, nsteps, rebuild_every, skin, target):
    """(loss, [dL/dmu, dL/dW_A, dL/dgamma], dL/dr) of the sphere steps; recs
    is advanced to the final state."""
    assert recs.dtype == RECORD_DTYPE and recs.flags.c_contiguous
    n = len(recs)
    fp = np.ascontiguousarray(fpar, np.float64)
    tg = np.ascontiguousarray(target, np.float64).reshape(-1)
    loss = np.zeros(1)
    grad = np.zeros(3)
    gr = np.zeros(n)
    rc = lib().orc_adjoint(recs.ctypes.data_as(C.c_void_p), n, C.byref(geom), _d(fp),
                           mobility, dt, nsteps, rebuild_every, skin, _d(tg), _d(loss),
                           _d(grad), _d(gr))
    if rc != 0:
        raise RuntimeError(f"orc_adjoint failed: {rc}")
    return loss[0], grad, gr


# ---- particle <-> mesh ------------------------------------------------------
# Meshes are numpy arrays indexed [comp, k, j, i] (x fastest in memory).
# Output meshes hold the valid cells; input meshes one ghost layer.
DEPOSIT_ONE_TO_ONE, DEPOSIT_TRILINEAR = 0, 1


def _mesh_shape(geom: Geom, ghost: int = 0):
    return tuple(int(geom.nbin[d]) + 2 * ghost for d in (2, 1, 0))


def deposit(recs, geom: Geom, scheme: int, comp0: int, ncomp: int):
    """ParticleToMesh of rdata[comp0:comp0+ncomp]; (mesh, particles outside)."""
    assert recs.dtype == RECORD_DTYPE and recs.flags.c_contiguous
    out = np.zeros((ncomp,) + _mesh_shape(geom))
    k = lib().orc_deposit(recs.ctypes.data_as(C.c_void_p), len(recs), C.byref(geom),
                          scheme, comp0, ncomp, _d(out))
    return out, k


def volume_fraction(recs, geom: Geom, scheme: int):
    assert recs.dtype == RECORD_DTYPE and recs.flags.c_contiguous
    out = np.zeros(_mesh_shape(geom))
    k = lib().orc_volume_fraction(recs.ctypes.data_as(C.c_void_p), len(recs),
                                  C.byref(geom), scheme, _d(out))
    return out, k


def count_per_cell(recs, geom: Geom):
    assert recs.dtype == RECORD_DTYPE and recs.flags.c_contiguous
    out = np.zeros(_mesh_shape(geom))
    k = lib().orc_count_per_cell(recs.ctypes.data_as(C.c_void_p), len(recs),
                                 C.byref(geom), _d(out))
    return out, k


def chem_update(grid_vol, npart, rdata, idata, mesh_vals, chempar, dt, nloop):
    """xferMeshToParticleAndUpdateChem of one particle on copies; returns
    (rdata, error bits)."""
    rd = np.array(rdata, np.float64)
    fm = np.array(mesh_vals, np.float64)
    ip = np.ascontiguousarray(idata, np.int32)
    cp = np.ascontiguousarray(chempar, np.float64)
    pv = rd[25:].copy()
    err = lib().orc_chem_update(float(grid_vol), float(npart), _d(rd), _d(fm), _d(pv),
                                float(dt), int(nloop), _d(cp), _i(ip))
    rd[25:] = pv
    return rd, err


def mesh_to_particle(recs, geom: Geom, X, vf, chempar, dt, nloop, cnc_scheme, vf_scheme):
    """bmx_calc_txfr_particle on recs (in place); X: (3, nz+2, ny+2, nx+2),
    vf: (nz+2, ny+2, nx+2).  Returns the error bits."""
    assert recs.dtype == RECORD_DTYPE and recs.flags.c_contiguous
    Xc = np.ascontiguousarray(X, np.float64)
    vc = np.ascontiguousarray(vf, np.float64)
    assert Xc.shape == (3,) + _mesh_shape(geom, 1) and vc.shape == _mesh_shape(geom, 1)
    cp = np.ascontiguousarray(chempar, np.float64)
    return lib().orc_mesh_to_particle(recs.ctypes.data_as(C.c_void_p), len(recs),
                                      C.byref(geom), _d(Xc), _d(vc), _d(cp), float(dt),
                                      int(nloop), int(cnc_scheme), int(vf_scheme))


def interp(x, field, geom: Geom, scheme: int):
    """Interpolation of a ghosted field (ncomp, nz+2, ny+2, nx+2) at x."""
    f = np.ascontiguousarray(field, np.float64)
    if f.ndim == 3:
        f = f[None]
    assert f.shape[1:] == _mesh_shape(geom, 1)
    xv = np.ascontiguousarray(x, np.float64)
    out = np.zeros(f.shape[0])
    lib().orc_interp(_d(xv), _d(f), C.byref(geom), int(scheme), f.shape[0], _d(out))
    return out


def pow_(x: float, y: float) -> float:
    return lib().orc_pow(float(x), float(y))


def pow_dd(x: float, y: float) -> float:
    """bmx_pow's general double-double path alone."""
    return lib().orc_pow_dd(float(x), float(y))


def pow_third(x: float, k: int) -> float:
    """bmx_pow's cube-root path for y = 1/3 (k = 1) or 2/3 (k = 2); 0 where
    it leaves the call to the general path."""
    return lib().orc_pow_third(float(x), int(k))


def split_particles(recs, geom: Geom, sp, seed: int, call: int, next_id: int, cpu: int):
    """split_particles on a copy of recs; returns the grown record array.
    sp = (max_vol, max_len, max_rad, brnch_prob, split_prob, overlap,
    split_len)."""
    assert recs.dtype == RECORD_DTYPE
    n = len(recs)
    out = np.zeros(3 * n + 1, dtype=RECORD_DTYPE)
    out[:n] = recs
    spa = np.ascontiguousarray(sp, np.float64)
    m = lib().orc_split_particles(out.ctypes.data_as(C.c_void_p), n, len(out),
                                  C.byref(geom), _d(spa), int(seed), int(call),
                                  int(next_id), int(cpu))
    if m < 0:
        raise RuntimeError("orc_split_particles failed")
    return out[:m].copy()


def fusion(recs, geom: Geom, fpar, mode: int, radius: float, fus, seed: int, call: int,
           next_id: int, cpu: int):
    """Tip fusion, interior fusion and cleanup on a copy of recs; returns
    (records, whether a tip fused).  fus = (fusion_probability,
    max_fusion_separation, bond_scaling_increment, max_seg_length)."""
    assert recs.dtype == RECORD_DTYPE
    n = len(recs)
    out = np.zeros(2 * n + 1, dtype=RECORD_DTYPE)
    out[:n] = recs
    fp = np.ascontiguousarray(fpar, np.float64)
    fu = np.ascontiguousarray(fus, np.float64)
    fused = C.c_int(0)
    m = lib().orc_fusion(out.ctypes.data_as(C.c_void_p), n, len(out), C.byref(geom),
                         _d(fp), int(mode), float(radius), _d(fu), int(seed), int(call),
                         int(next_id), int(cpu), C.byref(fused))
    if m < 0:
        raise RuntimeError(f"orc_fusion failed: {m}")
    return out[:m].copy(), bool(fused.value)
