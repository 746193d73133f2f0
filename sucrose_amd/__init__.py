"""sucrose_amd -- the SUCROSE per-step force-and-integrate path on one MI355X.

The product is ``libsucrose_hip.so`` (gfx950 HIP kernels behind the C ABI in
``include/sucrose.h``).  This module is a thin ctypes mirror of that ABI that
keeps the reference's routine names (``vv_initial_step``, ``init_nlist``,
``lj_calculate_forces`` ...), for Python callers and the tests.  It never
falls back to a CPU path: without the built library every entry point raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "libsucrose_hip.so")

# include/sucrose.h
FIELDS = [
    "RX", "RY", "RZ", "VX", "VY", "VZ", "FX", "FY", "FZ",
    "QW", "QX", "QY", "QZ", "LX", "LY", "LZ", "TX", "TY", "TZ",
    "MASS", "INV_MASS", "CHARGE", "SIG_X", "SIG_Y", "SIG_Z",
    "EPS_X", "EPS_Y", "EPS_Z", "IXX", "IYY", "IZZ",
    "IXX_INV", "IYY_INV", "IZZ_INV", "RX_OLD", "RY_OLD", "RZ_OLD",
]
FIELD = {f: k for k, f in enumerate(FIELDS)}
SCALARS = ["POT_ENG", "KIN_ENG", "KIN_ROT", "TEMP", "PRESSURE", "DENSITY", "VIRIAL",
           "VIRIAL_BONDED", "TARGET_TEMP", "DOF", "HB0", "HB1", "HB2", "ETA0", "ETA1", "ETA2",
           "PETA0", "PETA1", "PETA2"]
SCALAR = {s: k for k, s in enumerate(SCALARS)}
LJ, LJ12, LJW, WCA, MORSE, GB, GBR, ECP, ECPR, DH, GBW, GBC, GBWC, ECPW, ECPC, ECPWC, DIPOLE, COULOMB = range(1, 19)
ECP_BRENT_REFERENCE, ECP_CONVERGED, ECP_LEARNED = 0, 1, 2
REBUILD_REFERENCE, REBUILD_HALF_SKIN = 0, 1
INT_VV, INT_NHT, INT_LGV, INT_LGVC = 1, 4, 6, 7
REPORT_STATE_DATA, REPORT_TRAJECTORY, REPORT_CONFIG, REPORT_CHECKPOINT = 1, 2, 3, 4

ERRORS = {-1: "SX_ERR_ARG", -2: "SX_ERR_HIP", -3: "SX_ERR_STATE", -4: "SX_ERR_CAPACITY",
          -5: "SX_ERR_CONSTRAINT", -6: "SX_ERR_CUTOFF", -7: "SX_ERR_NOTEMP"}


class SucroseError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


_PD = C.POINTER(C.c_double)
_PI = C.POINTER(C.c_int)
_lib = None


class StepSpec(C.Structure):
    _fields_ = [("integrator", C.c_int), ("nforces", C.c_int), ("kinds", C.c_int * 4),
                ("params", (C.c_double * 8) * 4), ("bonded", C.c_int),
                ("dt", C.c_double), ("damping", C.c_double), ("first_step", C.c_int)]


def lib():
    """The loaded product library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            raise FileNotFoundError(f"{LIBPATH} is not built: run `python -m sucrose_amd.build`")
        L = C.CDLL(LIBPATH)
        L.sx_last_error.restype = C.c_char_p
        L.sx_create.argtypes = [C.c_int, C.POINTER(C.c_void_p)]
        vp = C.c_void_p
        for name, args in {
            "sx_destroy": [vp], "sx_synchronize": [vp],
            "sx_set_num_atoms": [vp, C.c_int, C.c_int],
            "sx_set_box": [vp, _PD, C.c_int], "sx_get_box": [vp, _PD],
            "sx_upload": [vp, C.c_int, _PD, C.c_int], "sx_download": [vp, C.c_int, _PD, C.c_int],
            "sx_field_ptr": [vp, C.c_int, C.POINTER(C.c_void_p)],
            "sx_set_scalar": [vp, C.c_int, C.c_double], "sx_get_scalar": [vp, C.c_int, _PD],
            "sx_set_constraints": [vp, C.c_int, _PI, _PI, _PD],
            "sx_initialize_system": [vp, C.c_int],
            "sx_set_velocity_to_temperature": [vp, C.c_double, C.c_int, C.c_int],
            "sx_initialize_forces": [vp], "sx_calculate_thermo": [vp],
            "sx_nlist_init": [vp, C.c_double, C.c_double], "sx_nlist_build": [vp],
            "sx_nlist_needs_rebuild": [vp, _PI], "sx_nlist_update": [vp],
            "sx_nlist_set_rule": [vp, C.c_int], "sx_nlist_rebuild_count": [vp, C.POINTER(C.c_long)],
            "sx_nlist_export": [vp, _PI, _PI, C.c_long, C.POINTER(C.c_long)],
            "sx_forces": [vp, C.c_int, _PD, C.c_int],
            "sx_pair_eval": [vp, C.c_int, _PD, C.c_int, C.c_long, _PI, _PI, _PD, _PD, _PD, _PD, _PD],
            "sx_set_ecp_solver": [vp, C.c_int], "sx_ecp_stats": [vp, C.POINTER(C.c_long)],
            "sx_set_pair_precision": [vp, C.c_int], "sx_get_pair_precision": [vp, C.POINTER(C.c_int)],
            "sx_nht_stats": [vp, C.POINTER(C.c_long)],
            "sx_set_bonded": [vp, C.c_int, C.c_int, _PI, C.c_double, C.c_double],
            "sx_bonded_forces": [vp],
            "sx_vv_initial_step": [vp, C.c_double], "sx_vv_final_step": [vp, C.c_double],
            "sx_vvc_initial_step": [vp, C.c_double], "sx_vvc_final_step": [vp, C.c_double],
            "sx_qq_initial_step": [vp, C.c_double], "sx_qq_final_step": [vp, C.c_double],
            "sx_nht_initial_step": [vp, C.c_double, _PD], "sx_nht_final_step": [vp, C.c_double, _PD],
            "sx_nhtc_initial_step": [vp, C.c_double, _PD], "sx_nhtc_final_step": [vp, C.c_double, _PD],
            "sx_lgv_initial_step": [vp, C.c_double, C.c_int, C.c_double],
            "sx_lgv_final_step": [vp, C.c_double, C.c_double],
            "sx_lgvc_initial_step": [vp, C.c_double, C.c_int, C.c_double],
            "sx_lgvc_final_step": [vp, C.c_double, C.c_double],
            "sx_apply_constraints_a": [vp, C.c_double], "sx_apply_constraints_b": [vp, C.c_double],
            "sx_run": [vp, C.POINTER(StepSpec), C.c_int],
            "sx_philox_normals": [vp, C.c_int, C.c_int, C.c_int, C.c_int, _PD],
            "sx_philox_normals_noalloc": [vp, C.c_int, C.c_int, C.c_int, C.c_int, _PD],
        }.items():
            getattr(L, name).argtypes = args
            getattr(L, name).restype = C.c_int
        L.sx_philox4x32_10.argtypes = [C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]
        L.sx_philox4x32_10.restype = None
        L.sx_profile_ecp.argtypes = [vp, C.c_int, _PD, C.c_int, C.c_int, _PD, _PD]
        L.sx_profile_ecp.restype = C.c_int
        L.sx_gen_ellipsoids_rsa.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, C.c_double,
                                            C.c_uint32, C.c_long, _PD, _PD]
        L.sx_gen_ellipsoids_rsa.restype = C.c_int
        L.sx_count_ellipsoid_overlaps.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, C.c_double,
                                                  _PD, _PD]
        L.sx_count_ellipsoid_overlaps.restype = C.c_long
        L.sx_gen_uniform_quaternions.argtypes = [C.c_int, C.c_uint32, _PD]
        L.sx_gen_uniform_quaternions.restype = None
        L.sx_gen_chains.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_uint32, C.c_long,
                                    _PD]
        L.sx_gen_chains.restype = C.c_int
        PF = C.POINTER(C.c_float)
        for name, args in {
            "sx_mlp_init": [vp, C.c_uint32], "sx_mlp_set_params": [vp, PF], "sx_mlp_get_params": [vp, PF],
            "sx_mlp_forward": [vp, C.c_long, PF, PF],
            "sx_mlp_loss_grad": [vp, C.c_long, PF, PF, PF, PF, PF],
            "sx_mlp_collect": [vp, C.c_int, _PD, C.c_int, C.c_long, C.POINTER(C.c_long)],
            "sx_mlp_samples": [vp, C.c_long, C.c_long, PF, PF], "sx_mlp_clear_samples": [vp],
            "sx_mlp_train": [vp, C.c_int, C.c_int, C.c_double, PF, C.c_uint32, PF],
            "sx_ecp_compare_solvers": [vp, C.c_int, _PD, C.c_int, C.c_int, _PD],
            "sx_batch_set": [vp, C.c_int, _PI, _PD, _PI, _PD, _PD],
            "sx_batch_run": [vp, C.c_int, C.c_double, C.c_double, C.c_int],
            "sx_batch_thermo": [vp, _PD],
            "sx_batch_autotune": [vp, C.c_double, C.c_double, C.c_int, C.c_int, _PI, _PD],
            "sx_batch_set_shape": [vp, C.c_int],
            "sx_set_atom_types": [vp, _PI], "sx_get_atom_types": [vp, _PI],
            "sx_set_density": [vp, C.c_double], "sx_generate_fcc_lattice": [vp, C.c_int, C.c_double, C.c_int],
            "sx_set_atom_values": [vp, C.c_int, C.c_int, C.c_int, _PD],
            "sx_set_target_pressure": [vp, C.c_double],
            "sx_copy_transform": [vp, C.c_int, _PI, _PI, _PD, C.c_int],
            "sx_transfer_forces": [vp, C.c_int, _PI, _PI],
            "sx_set_composites": [vp, C.c_int, _PI, _PI, _PD],
            "sx_composite_place": [vp], "sx_composite_transfer": [vp],
            "sx_init_reaction_field": [vp], "sx_eval_reaction_field": [vp],
            "sx_get_reaction_field": [vp, _PD, _PD],
            "sx_mc_barostat": [vp, C.POINTER(StepSpec), C.c_double, _PD, _PI],
            "sx_reporter_open": [vp, C.c_int, C.c_char_p, C.POINTER(C.c_void_p)],
            "sx_report": [vp, C.c_int],
            "sx_report_host": [vp, C.c_int, C.c_int, C.POINTER(_PD), _PI, _PD, _PD, C.c_int, _PI, _PI, _PD],
            "sx_reporter_set_config_index": [vp, C.c_int], "sx_reporter_flush": [vp],
            "sx_reporter_close": [vp],
            "sx_load_checkpoint": [vp, C.c_char_p, _PI],
            "sx_read_checkpoint_host": [C.c_char_p, _PI, _PI, _PD, _PI, C.c_int, C.c_int, C.POINTER(_PD), _PI,
                                        _PI, _PI, _PD],
        }.items():
            getattr(L, name).argtypes = args
            getattr(L, name).restype = C.c_int
        L.sx_mlp_num_params.restype = C.c_int
        L.sx_reporter_error.argtypes = [vp]
        L.sx_reporter_error.restype = C.c_char_p
        _lib = L
    return _lib


def philox4x32_10(ctr):
    """Host Philox4x32-10 with the reference's key (xmath.f90:42-76)."""
    c = (C.c_uint32 * 4)(*[int(x) & 0xFFFFFFFF for x in ctr])
    o = (C.c_uint32 * 4)()
    lib().sx_philox4x32_10(c, o)
    return np.array(list(o), dtype=np.uint32)


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _dp(a):
    return a.ctypes.data_as(_PD)


def gen_ellipsoids_rsa(n, semi_axes, L, seed, max_attempts=200_000_000):
    """Random sequential addition of identical ellipsoids; returns (pos (k,3), quat (k,4))."""
    pos = np.zeros(3 * n)
    quat = np.zeros(4 * n)
    a, b, c = semi_axes
    k = lib().sx_gen_ellipsoids_rsa(int(n), a, b, c, float(L), int(seed) & 0xFFFFFFFF, int(max_attempts),
                                    _dp(pos), _dp(quat))
    return pos.reshape(n, 3)[:k], quat.reshape(n, 4)[:k]


def uniform_quaternions(n, seed):
    q = np.zeros(4 * n)
    lib().sx_gen_uniform_quaternions(int(n), int(seed) & 0xFFFFFFFF, _dp(q))
    return q.reshape(n, 4)


def gen_chains(nchains, nbeads, L, bond=1.0, min_dist=0.8, seed=1, max_attempts=2_000_000_000):
    """Self-avoiding random-walk chains; returns positions (nchains*nbeads, 3)."""
    pos = np.zeros(3 * nchains * nbeads)
    k = lib().sx_gen_chains(int(nchains), int(nbeads), float(L), float(bond), float(min_dist),
                            int(seed) & 0xFFFFFFFF, int(max_attempts), _dp(pos))
    if k != nchains:
        raise RuntimeError(f"placed {k} of {nchains} chains")
    return pos.reshape(-1, 3)


def count_ellipsoid_overlaps(pos, quat, semi_axes, L):
    pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1)
    quat = np.ascontiguousarray(quat, dtype=np.float64).reshape(-1)
    a, b, c = semi_axes
    return lib().sx_count_ellipsoid_overlaps(len(pos) // 3, a, b, c, float(L), _dp(pos), _dp(quat))


class System:
    """One device-resident SUCROSE system (the reference's `module system`
    globals), driven with the reference's routine names."""

    def __init__(self, device: int = 0):
        self.L = lib()
        h = C.c_void_p()
        r = self.L.sx_create(device, C.byref(h))
        if r != 0:
            raise SucroseError(r, self.L.sx_last_error(None).decode())
        self.h = h
        self.n = 0

    def close(self):
        if self.h:
            self.L.sx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _ok(self, r):
        if r != 0:
            raise SucroseError(r, self.L.sx_last_error(self.h).decode())

    # ---- module system ----
    def set_num_atoms(self, num_atoms, rigid_atoms=1):
        self._ok(self.L.sx_set_num_atoms(self.h, int(num_atoms), int(rigid_atoms)))
        self.n = int(num_atoms) * int(rigid_atoms)

    def set_box(self, length, width=None, height=None, periodic=True):
        box = np.array([length, width if width is not None else length,
                        height if height is not None else length], dtype=np.float64)
        self._ok(self.L.sx_set_box(self.h, _dp(box), 1 if periodic else 0))

    def get_box(self):
        b = np.zeros(3)
        self._ok(self.L.sx_get_box(self.h, _dp(b)))
        return b

    def upload(self, field, values):
        a = np.ascontiguousarray(values, dtype=np.float64)
        if a.shape != (self.n,):
            a = np.ascontiguousarray(np.broadcast_to(a, (self.n,)), dtype=np.float64)
        self._ok(self.L.sx_upload(self.h, FIELD[field.upper()], _dp(a), self.n))

    def download(self, field):
        a = np.zeros(self.n)
        self._ok(self.L.sx_download(self.h, FIELD[field.upper()], _dp(a), self.n))
        return a

    def field_ptr(self, field):
        p = C.c_void_p()
        self._ok(self.L.sx_field_ptr(self.h, FIELD[field.upper()], C.byref(p)))
        return p.value

    def set_scalar(self, name, v):
        self._ok(self.L.sx_set_scalar(self.h, SCALAR[name.upper()], float(v)))

    def get_scalar(self, name):
        v = C.c_double()
        self._ok(self.L.sx_get_scalar(self.h, SCALAR[name.upper()], C.byref(v)))
        return v.value

    def set_target_temperature(self, t):
        self.set_scalar("TARGET_TEMP", t)

    def set_constraints(self, bbi1, bbj1, lengths):
        bi = np.ascontiguousarray(bbi1, dtype=np.int32)
        bj = np.ascontiguousarray(bbj1, dtype=np.int32)
        bl = np.ascontiguousarray(lengths, dtype=np.float64)
        self._ok(self.L.sx_set_constraints(self.h, len(bi), bi.ctypes.data_as(_PI), bj.ctypes.data_as(_PI), _dp(bl)))

    def initialize_system(self, remove_com=True):
        self._ok(self.L.sx_initialize_system(self.h, 1 if remove_com else 0))

    def set_velocity_to_temperature(self, t, seed, remove_com=True):
        self._ok(self.L.sx_set_velocity_to_temperature(self.h, float(t), int(seed), 1 if remove_com else 0))

    def initialize_forces(self):
        self._ok(self.L.sx_initialize_forces(self.h))

    def calculate_thermo(self):
        self._ok(self.L.sx_calculate_thermo(self.h))

    def synchronize(self):
        self._ok(self.L.sx_synchronize(self.h))

    # ---- module neighbour_list ----
    def init_nlist(self, skin, cutoff):
        self._ok(self.L.sx_nlist_init(self.h, float(skin), float(cutoff)))

    def build_nlist(self):
        self._ok(self.L.sx_nlist_build(self.h))

    def needs_rebuild(self):
        o = C.c_int()
        self._ok(self.L.sx_nlist_needs_rebuild(self.h, C.byref(o)))
        return bool(o.value)

    def update_nlist(self):
        self._ok(self.L.sx_nlist_update(self.h))

    def set_rebuild_rule(self, rule):
        self._ok(self.L.sx_nlist_set_rule(self.h, int(rule)))

    def rebuild_count(self):
        o = C.c_long()
        self._ok(self.L.sx_nlist_rebuild_count(self.h, C.byref(o)))
        return o.value

    def export_nlist(self):
        """(offset, nlist) in the reference layout: 1-based, offset(N) = nnz+1."""
        nnz = C.c_long()
        off = np.zeros(self.n, np.int32)
        self._ok(self.L.sx_nlist_export(self.h, off.ctypes.data_as(_PI), None, 0, C.byref(nnz)))
        nb = np.zeros(max(nnz.value, 1), np.int32)
        self._ok(self.L.sx_nlist_export(self.h, off.ctypes.data_as(_PI), nb.ctypes.data_as(_PI),
                                        len(nb), C.byref(nnz)))
        return off, nb[:nnz.value]

    # ---- forces ----
    def forces(self, kind, *params):
        p = np.array(params if params else [0.0], dtype=np.float64)
        self._ok(self.L.sx_forces(self.h, int(kind), _dp(p), len(params)))

    def pair_eval(self, kind, i1, j1, *params):
        """Per-pair `*_calculate_forces(I, J, ...)` for 1-based index arrays."""
        i1 = np.ascontiguousarray(i1, dtype=np.int32)
        j1 = np.ascontiguousarray(j1, dtype=np.int32)
        npairs = len(i1)
        p = np.array(params if params else [0.0], dtype=np.float64)
        pe = np.zeros(npairs); f = np.zeros((npairs, 3)); ti = np.zeros((npairs, 3))
        tj = np.zeros((npairs, 3)); lam = np.zeros(npairs)
        self._ok(self.L.sx_pair_eval(self.h, int(kind), _dp(p), len(params), npairs, i1.ctypes.data_as(_PI),
                                     j1.ctypes.data_as(_PI), _dp(pe), _dp(f), _dp(ti), _dp(tj), _dp(lam)))
        return pe, f, ti, tj, lam

    def set_ecp_solver(self, solver):
        self._ok(self.L.sx_set_ecp_solver(self.h, int(solver)))

    def set_pair_precision(self, bits):
        """32: FP32 pair terms for LJ, WCA and Debye-Hueckel, and GB (variant 0)
        in its form with fewer divisions; 64: the reference's FP64 expressions
        throughout (the default)."""
        self._ok(self.L.sx_set_pair_precision(self.h, int(bits)))

    def pair_precision(self):
        b = C.c_int()
        self._ok(self.L.sx_get_pair_precision(self.h, C.byref(b)))
        return b.value

    def ecp_stats(self):
        o = (C.c_long * 4)()
        self._ok(self.L.sx_ecp_stats(self.h, o))
        return list(o)

    def nht_stats(self):
        """Cumulative Nose-Hoover (parallel evaluation) counts: Newton
        iterations, calls, calls evaluated window by window because their
        passes did not converge."""
        o = (C.c_long * 3)()
        self._ok(self.L.sx_nht_stats(self.h, o))
        return list(o)

    def profile_ecp(self, kind, params=(), nrep=10):
        """HIP-event timings (ms) of the ECP pipeline stages and per-run counts."""
        p = np.array(params if params else [0.0], dtype=np.float64)
        ms = np.zeros(4)
        stats = np.zeros(3)
        self._ok(self.L.sx_profile_ecp(self.h, int(kind), _dp(p), len(params), int(nrep), _dp(ms), _dp(stats)))
        return ms, stats

    # ---- learned lambda* warm start ----
    def mlp_init(self, seed=1):
        self._ok(self.L.sx_mlp_init(self.h, int(seed) & 0xFFFFFFFF))

    def mlp_set_params(self, p):
        a = np.ascontiguousarray(p, dtype=np.float32)
        if a.size != self.L.sx_mlp_num_params():
            raise ValueError("wrong parameter count")
        self._ok(self.L.sx_mlp_set_params(self.h, _fp(a)))

    def mlp_get_params(self):
        a = np.zeros(self.L.sx_mlp_num_params(), dtype=np.float32)
        self._ok(self.L.sx_mlp_get_params(self.h, _fp(a)))
        return a

    def mlp_forward(self, X):
        X = np.ascontiguousarray(X, dtype=np.float32).reshape(-1, 7)
        Y = np.zeros((len(X), 2), dtype=np.float32)
        self._ok(self.L.sx_mlp_forward(self.h, len(X), _fp(X), _fp(Y)))
        return Y

    def mlp_loss_grad(self, X, T, w=(1.0, 1.0)):
        X = np.ascontiguousarray(X, dtype=np.float32).reshape(-1, 7)
        T = np.ascontiguousarray(T, dtype=np.float32).reshape(-1, 2)
        wv = np.array(w, dtype=np.float32)
        loss = np.zeros(1, dtype=np.float32)
        g = np.zeros(self.L.sx_mlp_num_params(), dtype=np.float32)
        self._ok(self.L.sx_mlp_loss_grad(self.h, len(X), _fp(X), _fp(T), _fp(wv), _fp(loss), _fp(g)))
        return float(loss[0]), g

    def mlp_collect(self, kind, params=(), capacity=1 << 21):
        p = np.array(params if params else [0.0], dtype=np.float64)
        tot = C.c_long()
        self._ok(self.L.sx_mlp_collect(self.h, int(kind), _dp(p), len(params), int(capacity), C.byref(tot)))
        return tot.value

    def mlp_samples(self, first, n):
        X = np.zeros((n, 7), dtype=np.float32)
        T = np.zeros((n, 2), dtype=np.float32)
        self._ok(self.L.sx_mlp_samples(self.h, int(first), int(n), _fp(X), _fp(T)))
        return X, T

    def mlp_clear_samples(self):
        self._ok(self.L.sx_mlp_clear_samples(self.h))

    def mlp_train(self, steps, batch=4096, lr=1e-3, w=(1.0, 1.0), seed=7):
        wv = np.array(w, dtype=np.float32)
        trace = np.zeros(steps, dtype=np.float32)
        self._ok(self.L.sx_mlp_train(self.h, int(steps), int(batch), float(lr), _fp(wv), int(seed) & 0xFFFFFFFF,
                                     _fp(trace)))
        return trace

    def compare_solvers(self, kind, params=(), nrep=10):
        """Contact-parameter solvers on the current ECP work list (times in ms)."""
        p = np.array(params if params else [0.0], dtype=np.float64)
        o = np.zeros(19)
        self._ok(self.L.sx_ecp_compare_solvers(self.h, int(kind), _dp(p), len(params), int(nrep), _dp(o)))
        keys = ["pairs", "ms_brent", "ms_newton_plain", "ms_network", "ms_newton_warm", "ms_solve_plain",
                "ms_solve_learned", "ms_brent_tight", "iters_plain", "iters_warm", "wave_iters_plain",
                "wave_iters_warm", "lambda0_err_mean", "lambda0_err_max", "lambda0_err_rms",
                "lambda_warm_vs_plain_max", "iters_brent_tight", "wave_iters_brent_tight", "proven_out_of_range"]
        return dict(zip(keys, o.tolist()))

    # ---- batched replicas ----
    # (particles per force workgroup, lanes per particle)
    BATCH_SHAPES = ((64, 1), (16, 4), (32, 4), (32, 8), (64, 8), (16, 64), (32, 64), (64, 64), (0, 256), (0, 512))

    def batch_set(self, counts, boxes, kinds, params, temps):
        """Splits the particles into len(counts) independent systems (contiguous
        ranges); params is (nsys, <=4) per system (LJ: sigma, epsilon, cutoff;
        GB: cutoff, mu, nu)."""
        cnt = np.ascontiguousarray(counts, dtype=np.int32)
        ns = len(cnt)
        bx = np.ascontiguousarray(boxes, dtype=np.float64).reshape(ns)
        kd = np.ascontiguousarray(kinds, dtype=np.int32).reshape(ns)
        pr = np.zeros((ns, 4))
        pa = np.asarray(params, dtype=np.float64).reshape(ns, -1)
        pr[:, :pa.shape[1]] = pa
        tp = np.ascontiguousarray(temps, dtype=np.float64).reshape(ns)
        self._ok(self.L.sx_batch_set(self.h, ns, cnt.ctypes.data_as(_PI), _dp(bx), kd.ctypes.data_as(_PI),
                                     _dp(pr), _dp(tp)))
        self._batch_n = ns

    def batch_run(self, nsteps, dt, damping=1.0, first_step=1):
        self._ok(self.L.sx_batch_run(self.h, int(nsteps), float(dt), float(damping), int(first_step)))

    def batch_thermo(self):
        """(nsys, 4): KE, KErot, PE, T of each system after the last step."""
        out = np.zeros(4 * self._batch_n)
        self._ok(self.L.sx_batch_thermo(self.h, _dp(out)))
        return out.reshape(-1, 4)

    def batch_autotune(self, dt, damping=1.0, first_step=1, nrep=5):
        """Times every launch shape from the current state (restored after);
        returns (chosen index, ms per step of each shape)."""
        ch = C.c_int(0)
        ms = np.zeros(len(self.BATCH_SHAPES))
        self._ok(self.L.sx_batch_autotune(self.h, float(dt), float(damping), int(first_step), int(nrep),
                                          C.byref(ch), _dp(ms)))
        return ch.value, ms

    def batch_set_shape(self, choice):
        self._ok(self.L.sx_batch_set_shape(self.h, int(choice)))

    def set_bonded(self, kind, idx1, coeff, angle=0.0):
        a = np.ascontiguousarray(idx1, dtype=np.int32).reshape(-1)
        npos = 3 if kind in (1, 2, 5) else 4
        self._ok(self.L.sx_set_bonded(self.h, int(kind), len(a) // npos, a.ctypes.data_as(_PI),
                                      float(coeff), float(angle)))

    def bonded_forces(self):
        self._ok(self.L.sx_bonded_forces(self.h))

    # ---- integrators ----
    def vv_initial_step(self, dt): self._ok(self.L.sx_vv_initial_step(self.h, dt))
    def vv_final_step(self, dt): self._ok(self.L.sx_vv_final_step(self.h, dt))
    def vvc_initial_step(self, dt): self._ok(self.L.sx_vvc_initial_step(self.h, dt))
    def vvc_final_step(self, dt): self._ok(self.L.sx_vvc_final_step(self.h, dt))
    def qq_initial_step(self, dt): self._ok(self.L.sx_qq_initial_step(self.h, dt))
    def qq_final_step(self, dt): self._ok(self.L.sx_qq_final_step(self.h, dt))

    def _cpl(self, coupling):
        return C.byref(C.c_double(coupling)) if coupling is not None else None

    def nht_initial_step(self, dt, coupling=None): self._ok(self.L.sx_nht_initial_step(self.h, dt, self._cpl(coupling)))
    def nht_final_step(self, dt, coupling=None): self._ok(self.L.sx_nht_final_step(self.h, dt, self._cpl(coupling)))
    def nhtc_initial_step(self, dt, coupling=None): self._ok(self.L.sx_nhtc_initial_step(self.h, dt, self._cpl(coupling)))
    def nhtc_final_step(self, dt, coupling=None): self._ok(self.L.sx_nhtc_final_step(self.h, dt, self._cpl(coupling)))
    def lgv_initial_step(self, dt, step, damping=1.0): self._ok(self.L.sx_lgv_initial_step(self.h, dt, int(step), damping))
    def lgv_final_step(self, dt, damping=1.0): self._ok(self.L.sx_lgv_final_step(self.h, dt, damping))
    def lgvc_initial_step(self, dt, step, damping=1.0): self._ok(self.L.sx_lgvc_initial_step(self.h, dt, int(step), damping))
    def lgvc_final_step(self, dt, damping=1.0): self._ok(self.L.sx_lgvc_final_step(self.h, dt, damping))
    def apply_constraints_a(self, dt): self._ok(self.L.sx_apply_constraints_a(self.h, dt))
    def apply_constraints_b(self, dt): self._ok(self.L.sx_apply_constraints_b(self.h, dt))

    def run(self, integrator, forces, nsteps, dt, damping=1.0, first_step=1, bonded=False):
        """nsteps whole steps replayed from one captured HIP graph.
        forces: list of (kind, params...) tuples."""
        sp = StepSpec()
        sp.integrator = integrator
        sp.nforces = len(forces)
        for k, fp in enumerate(forces):
            sp.kinds[k] = fp[0]
            for q, v in enumerate(fp[1:]):
                sp.params[k][q] = v
        sp.bonded = 1 if bonded else 0
        sp.dt = dt
        sp.damping = damping
        sp.first_step = first_step
        self._ok(self.L.sx_run(self.h, C.byref(sp), int(nsteps)))

    def philox_normals(self, n, c2, c3, c4):
        out = np.zeros(3 * n)
        self._ok(self.L.sx_philox_normals(self.h, n, c2, c3, c4, _dp(out)))
        return out.reshape(n, 3)

    def philox_normals_noalloc(self, n, c2, c3, c4):
        """philox_normals staged in the GPU's fixed scratch (no allocation)."""
        out = np.zeros(3 * n)
        self._ok(self.L.sx_philox_normals_noalloc(self.h, n, c2, c3, c4, _dp(out)))
        return out.reshape(n, 3)

    # ---- whole-state helpers ----
    def load_arrays(self, arrays: dict):
        for k, v in arrays.items():
            self.upload(k, v)

    def state(self, fields=FIELDS):
        return {f: self.download(f) for f in fields}

    # ---- composite rigid bodies (system.f90:743-881) ----
    def copy_transform(self, from1, to1, offsets, set_inertia=None):
        a = np.ascontiguousarray(np.atleast_1d(from1), dtype=np.int32)
        b = np.ascontiguousarray(np.atleast_1d(to1), dtype=np.int32)
        off = np.ascontiguousarray(np.reshape(offsets, (-1, 3)), dtype=np.float64)
        si = -1 if set_inertia is None else int(bool(set_inertia))
        self._ok(self.L.sx_copy_transform(self.h, len(a), a.ctypes.data_as(_PI), b.ctypes.data_as(_PI), _dp(off), si))

    def transfer_forces(self, from1, to1):
        a = np.ascontiguousarray(np.atleast_1d(from1), dtype=np.int32)
        b = np.ascontiguousarray(np.atleast_1d(to1), dtype=np.int32)
        self._ok(self.L.sx_transfer_forces(self.h, len(a), a.ctypes.data_as(_PI), b.ctypes.data_as(_PI)))

    def set_composites(self, centre1, sat1, offsets):
        a = np.ascontiguousarray(centre1, dtype=np.int32)
        b = np.ascontiguousarray(sat1, dtype=np.int32)
        off = np.ascontiguousarray(np.reshape(offsets, (-1, 3)), dtype=np.float64)
        self._ok(self.L.sx_set_composites(self.h, len(a), a.ctypes.data_as(_PI), b.ctypes.data_as(_PI), _dp(off)))

    def composite_place(self):
        self._ok(self.L.sx_composite_place(self.h))

    def composite_transfer(self):
        self._ok(self.L.sx_composite_transfer(self.h))

    # ---- reaction field of the point dipoles (module dipole) ----
    def init_reaction_field(self):
        self._ok(self.L.sx_init_reaction_field(self.h))

    def eval_reaction_field(self):
        self._ok(self.L.sx_eval_reaction_field(self.h))

    def get_reaction_field(self):
        """(DIP as a (3, n) array, RFL as a 3-vector shared by all particles)."""
        dip = np.zeros(3 * self.n)
        rfl = np.zeros(3)
        self._ok(self.L.sx_get_reaction_field(self.h, _dp(dip), _dp(rfl)))
        return dip.reshape(3, self.n), rfl

    # ---- Monte Carlo barostat ----
    def set_target_pressure(self, p):
        self._ok(self.L.sx_set_target_pressure(self.h, float(p)))

    def mc_barostat(self, forces, u, volume_change=None, bonded=False):
        """mc_barostat (integrators.f90:864-945): compute_potential is the
        list rebuilt at the current box plus the pair loops `forces` (list of
        (kind, params...)) and the bonded terms; u are the move's two uniform
        numbers.  Returns True when the volume move is accepted."""
        sp = StepSpec()
        sp.nforces = len(forces)
        for k, fp in enumerate(forces):
            sp.kinds[k] = fp[0]
            for q, v in enumerate(fp[1:]):
                sp.params[k][q] = v
        sp.bonded = 1 if bonded else 0
        uu = np.ascontiguousarray(u, dtype=np.float64)
        acc = C.c_int()
        dv = -1.0 if volume_change is None else float(volume_change)
        self._ok(self.L.sx_mc_barostat(self.h, C.byref(sp), dv, _dp(uu), C.byref(acc)))
        return bool(acc.value)

    # ---- set-up routines of module system (system.f90:43-360) ----
    def set_density(self, dens):
        self._ok(self.L.sx_set_density(self.h, float(dens)))

    def set_periodic_boundary(self):
        b = self.get_box()
        self._ok(self.L.sx_set_box(self.h, _dp(b), 1))

    def generate_fcc_lattice(self, unit_cells, dens, rigid_atoms=None):
        self._ok(self.L.sx_generate_fcc_lattice(self.h, int(unit_cells), float(dens), int(rigid_atoms or 0)))
        if self.n == 0:
            self.n = 4 * int(unit_cells) ** 3 * int(rigid_atoms or 1)

    def _set_atom(self, fields, id1, values):
        for f, v in zip(fields, values):
            a = np.array([float(v)])
            self._ok(self.L.sx_set_atom_values(self.h, FIELD[f], int(id1), 1, _dp(a)))

    def set_atom_mass(self, id1, m):
        self._set_atom(["MASS"], id1, [m])

    def set_atom_charge(self, id1, q):
        self._set_atom(["CHARGE"], id1, [q])

    def set_atom_shape(self, id1, length, width, height):
        if length == 0.0 or width == 0.0 or height == 0.0:
            raise SucroseError(-1, "Error: Shape dimensions cannot be zero.")
        self._set_atom(["SIG_X", "SIG_Y", "SIG_Z"], id1, [length, width, height])

    def set_atom_energy(self, id1, ex, ey, ez):
        self._set_atom(["EPS_X", "EPS_Y", "EPS_Z"], id1, [ex, ey, ez])

    def set_atom_position(self, id1, x, y, z):
        self._set_atom(["RX", "RY", "RZ"], id1, [x, y, z])

    def set_atom_velocity(self, id1, vx, vy, vz):
        self._set_atom(["VX", "VY", "VZ"], id1, [vx, vy, vz])

    def set_atom_inertia(self, id1, ix, iy, iz):
        self._set_atom(["IXX", "IYY", "IZZ"], id1, [ix, iy, iz])

    def set_atom_type(self, id1, t):
        if not 1 <= int(id1) <= self.n:
            raise SucroseError(-1, f"Error: Atom index {id1} does not exist in system of {self.n} atoms.")
        rt = self.get_atom_types()
        rt[int(id1) - 1] = int(t)
        self.set_atom_types(rt)

    # ---- atom types and checkpoints ----
    def set_atom_types(self, rtype):
        a = np.ascontiguousarray(rtype, dtype=np.int32)
        if a.shape != (self.n,):
            raise ValueError("one type per atom")
        self._ok(self.L.sx_set_atom_types(self.h, a.ctypes.data_as(_PI)))

    def get_atom_types(self):
        a = np.zeros(self.n, np.int32)
        self._ok(self.L.sx_get_atom_types(self.h, a.ctypes.data_as(_PI)))
        return a

    def load_checkpoint(self, filename):
        """load_checkpoint (reporters.f90:188-221); returns config_index."""
        ci = C.c_int()
        n = C.c_int()
        self.L.sx_read_checkpoint_host(os.fsencode(filename), C.byref(n), None, None, None, 0, 0, None, None, None,
                                       None, None)
        self._ok(self.L.sx_load_checkpoint(self.h, os.fsencode(filename), C.byref(ci)))
        self.n = n.value
        return ci.value

    def reporter(self, kind, filename):
        """A reporter of this system's state (reporters.f90)."""
        return Reporter(kind, filename, self)


class Reporter:
    """One open reporter file: state data, trajectory, configuration or
    checkpoint frames in the reference's text formats."""

    def __init__(self, kind, filename, system=None):
        self.L = lib()
        self.kind = int(kind)
        h = C.c_void_p()
        ctx = system.h if system is not None else None
        r = self.L.sx_reporter_open(ctx, self.kind, os.fsencode(filename), C.byref(h))
        if r != 0:
            msg = self.L.sx_last_error(ctx).decode() if ctx else "cannot open reporter"
            raise SucroseError(r, msg)
        self.h = h
        self.system = system

    def _ok(self, r):
        if r != 0:
            raise SucroseError(r, self.L.sx_reporter_error(self.h).decode())

    def report(self, step=0):
        self._ok(self.L.sx_report(self.h, int(step)))

    def report_host(self, step, fields: dict, scal, box, rtype=None, bonds=None):
        """A frame from host arrays: fields maps FIELDS names to arrays;
        scal is the SX_NSCALARS vector; bonds = (bbi1, bbj1, length)."""
        n = len(next(iter(fields.values())))
        keep = {FIELD[k.upper()]: np.ascontiguousarray(v, dtype=np.float64) for k, v in fields.items()}
        ptrs = (_PD * len(FIELDS))()
        for k, v in keep.items():
            ptrs[k] = _dp(v)
        sc = np.ascontiguousarray(scal, dtype=np.float64)
        bx = np.ascontiguousarray(box, dtype=np.float64)
        rt = np.ascontiguousarray(rtype if rtype is not None else np.full(n, -1), dtype=np.int32)
        if bonds is not None and len(bonds[0]):
            bi = np.ascontiguousarray(bonds[0], dtype=np.int32)
            bj = np.ascontiguousarray(bonds[1], dtype=np.int32)
            bl = np.ascontiguousarray(bonds[2], dtype=np.float64)
            nc, pbi, pbj, pbl = len(bi), bi.ctypes.data_as(_PI), bj.ctypes.data_as(_PI), _dp(bl)
        else:
            nc, pbi, pbj, pbl = 0, None, None, None
        self._ok(self.L.sx_report_host(self.h, int(step), n, ptrs, rt.ctypes.data_as(_PI), _dp(sc), _dp(bx), nc,
                                       pbi, pbj, pbl))

    def set_config_index(self, k):
        self._ok(self.L.sx_reporter_set_config_index(self.h, int(k)))

    def flush(self):
        self._ok(self.L.sx_reporter_flush(self.h))

    def close(self):
        if self.h:
            r = self.L.sx_reporter_close(self.h)
            self.h = None
            if r != 0:
                raise SucroseError(r, "reporter close failed")

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def read_checkpoint(filename):
    """The host parse of load_checkpoint: dict with n, nc, box, config_index,
    fields (FIELDS name -> array, the checkpoint's columns and the defaults
    of set_num_atoms elsewhere), rtype and the 1-based bonds."""
    L = lib()
    n, nc, ci = C.c_int(), C.c_int(), C.c_int()
    box = np.zeros(3)
    r = L.sx_read_checkpoint_host(os.fsencode(filename), C.byref(n), C.byref(nc), _dp(box), C.byref(ci), 0, 0,
                                  None, None, None, None, None)
    if r not in (0, -4):
        raise SucroseError(r, f"cannot read checkpoint {filename}")
    N, NC = n.value, nc.value
    arrays = {f: np.zeros(N) for f in FIELDS}
    ptrs = (_PD * len(FIELDS))()
    for k, f in enumerate(FIELDS):
        ptrs[k] = _dp(arrays[f])
    rt = np.zeros(N, np.int32)
    bi = np.zeros(max(NC, 1), np.int32); bj = np.zeros(max(NC, 1), np.int32); bl = np.zeros(max(NC, 1))
    r = L.sx_read_checkpoint_host(os.fsencode(filename), C.byref(n), C.byref(nc), _dp(box), C.byref(ci), N, NC, ptrs,
                                  rt.ctypes.data_as(_PI), bi.ctypes.data_as(_PI), bj.ctypes.data_as(_PI), _dp(bl))
    if r != 0:
        raise SucroseError(r, f"cannot read checkpoint {filename}")
    return {"n": N, "nc": NC, "box": box, "config_index": ci.value, "fields": arrays, "rtype": rt,
            "bonds": (bi[:NC], bj[:NC], bl[:NC])}
