"""ctypes front end of the CPU checker -- TEST INFRASTRUCTURE ONLY.

Two implementations of the same reference routines, driven on identical FP64
state:

* ``Oracle``    -- the C restatement in ``oracle/oracle.c`` (``liboracle.so``).
* ``Reference`` -- the SUCROSE reference itself, compiled from its Fortran
  sources by ``oracle/Makefile`` into ``oracle/_ref/libsucrose_ref.so``
  (present only where that build ran).

Only the tests, the smoke check and the benchmark's CPU-baseline leg import
this module.  The product (``sucrose_amd``) never does.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

# Per-particle fields in the order of the module arrays of `module system`
# (fortran/system.f90:7-22) plus RX_old.. ; shared with ref_harness.f90.
FIELDS = [
    "rx", "ry", "rz", "vx", "vy", "vz", "fx", "fy", "fz",
    "qw", "qx", "qy", "qz", "lx", "ly", "lz", "tx", "ty", "tz",
    "mass", "inv_mass", "charge", "sig_x", "sig_y", "sig_z",
    "eps_x", "eps_y", "eps_z", "ixx", "iyy", "izz",
    "ixx_inv", "iyy_inv", "izz_inv", "rx_old", "ry_old", "rz_old",
]
FIDX = {f: k for k, f in enumerate(FIELDS)}
SCALARS = [
    "pot_eng", "kin_eng", "kin_rot", "temp", "pressure", "density", "virial",
    "virial_bonded", "target_temp", "dof",
    "hb0", "hb1", "hb2", "eta0", "eta1", "eta2", "peta0", "peta1", "peta2",
]
SIDX = {s: k for k, s in enumerate(SCALARS)}

# force kinds (include/sucrose.h SX_KIND_*)
LJ, LJ12, LJW, WCA, MORSE, GB, GBR, ECP, ECPR, DH, GBW, GBC, GBWC, ECPW, ECPC, ECPWC, DIPOLE, COULOMB = range(1, 19)
# integrator kinds used by ref_integrate
VV, VVC, QQ, NHT, NHTC, LGV, LGVC = range(1, 8)


@dataclass
class State:
    """Host copy of one SUCROSE system in FP64 (reference layout, 0-based)."""
    n: int
    box: np.ndarray
    periodic: bool = True
    nrigid: int = 1
    buf: np.ndarray = None          # (n, 37) Fortran order
    scal: np.ndarray = None         # (19,)
    bbi: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int32))
    bbj: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int32))
    bb_len: np.ndarray = field(default_factory=lambda: np.zeros(0))

    def __post_init__(self):
        self.box = np.asarray(self.box, dtype=np.float64).reshape(3)
        if self.buf is None:
            self.buf = np.zeros((self.n, len(FIELDS)), order="F")
            self.buf[:, FIDX["qw"]] = 1.0
        if self.scal is None:
            self.scal = np.zeros(len(SCALARS))
            self.scal[SIDX["hb0"]:SIDX["hb2"] + 1] = 1.0
            self.scal[SIDX["density"]] = self.n / self.nrigid / float(np.prod(self.box))

    def __getitem__(self, name):
        if name in FIDX:
            return self.buf[:, FIDX[name]]
        return self.scal[SIDX[name]]

    def __setitem__(self, name, value):
        if name in FIDX:
            self.buf[:, FIDX[name]] = value
        else:
            self.scal[SIDX[name]] = value

    @property
    def nc(self):
        return len(self.bbi)

    def copy(self):
        return State(self.n, self.box.copy(), self.periodic, self.nrigid,
                     np.asfortranarray(self.buf.copy()), self.scal.copy(),
                     self.bbi.copy(), self.bbj.copy(), self.bb_len.copy())


# ---------------------------------------------------------------------------
# C restatement
# ---------------------------------------------------------------------------
_PD = C.POINTER(C.c_double)
_PI = C.POINTER(C.c_int)
_PL = C.POINTER(C.c_long)


class OrSys(C.Structure):
    _fields_ = ([("n", C.c_int), ("nc", C.c_int), ("nrigid", C.c_int), ("dof", C.c_int),
                 ("periodic", C.c_int), ("_pad", C.c_int), ("box", C.c_double * 3)]
                + [(f, _PD) for f in FIELDS]
                + [("bbi", _PI), ("bbj", _PI), ("bb_len", _PD)]
                + [(s, C.c_double) for s in ("pot_eng", "kin_eng", "kin_rot", "temp", "pressure",
                                             "density", "virial", "virial_bonded", "target_temp")]
                + [("hb", C.c_double * 3), ("eta", C.c_double * 3), ("p_eta", C.c_double * 3)])


class OrNlist(C.Structure):
    _fields_ = [("n", C.c_int), ("cap", C.c_long), ("nnz", C.c_long),
                ("offset", _PL), ("nbr", _PI),
                ("x_last", _PD), ("y_last", _PD), ("z_last", _PD),
                ("cutoff", C.c_double), ("skin", C.c_double), ("r_skin_sq", C.c_double)]


def _dp(a):
    return a.ctypes.data_as(_PD)


class Oracle:
    """The C restatement (liboracle.so)."""

    def __init__(self, path=None):
        path = path or os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} not built (make -C oracle oracle)")
        self.lib = C.CDLL(path)
        L = self.lib
        L.or_philox4x32_10.argtypes = [C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]
        L.or_philox_normal_3seq.argtypes = [C.c_int32] * 4 + [_PD] * 3
        L.or_cholesky_solve.argtypes = [_PD, _PD, _PD]
        L.or_ecp_brent.argtypes = [_PD, _PD, _PD, _PI]
        L.or_ecp_brent.restype = C.c_double
        L.or_ecp_optim_dist.argtypes = [C.c_double, _PD, _PD, _PD]
        L.or_ecp_optim_dist.restype = C.c_double
        L.or_exprel.argtypes = [C.c_double]
        L.or_exprel.restype = C.c_double
        L.or_nlist_build.restype = C.c_int
        L.or_nlist_needs_rebuild.restype = C.c_int
        L.or_apply_constraints_a.restype = C.c_int
        L.or_apply_constraints_b.restype = C.c_int

    # -- struct plumbing --
    def _sys(self, st: State):
        s = OrSys()
        s.n, s.nc, s.nrigid = st.n, st.nc, st.nrigid
        s.dof = int(st["dof"])
        s.periodic = 1 if st.periodic else 0
        s.box[:] = st.box
        assert st.buf.flags.f_contiguous
        for k, f in enumerate(FIELDS):
            setattr(s, f, C.cast(st.buf.ctypes.data + k * st.n * 8, _PD))
        st._bbi32 = np.ascontiguousarray(st.bbi, dtype=np.int32)
        st._bbj32 = np.ascontiguousarray(st.bbj, dtype=np.int32)
        st.bb_len = np.ascontiguousarray(st.bb_len, dtype=np.float64)
        s.bbi = st._bbi32.ctypes.data_as(_PI)
        s.bbj = st._bbj32.ctypes.data_as(_PI)
        s.bb_len = _dp(st.bb_len)
        for name in ("pot_eng", "kin_eng", "kin_rot", "temp", "pressure", "density",
                     "virial", "virial_bonded", "target_temp"):
            setattr(s, name, float(st[name]))
        s.hb[:] = st.scal[SIDX["hb0"]:SIDX["hb2"] + 1]
        s.eta[:] = st.scal[SIDX["eta0"]:SIDX["eta2"] + 1]
        s.p_eta[:] = st.scal[SIDX["peta0"]:SIDX["peta2"] + 1]
        return s

    def _back(self, s: OrSys, st: State):
        st.box[:] = list(s.box)
        for name in ("pot_eng", "kin_eng", "kin_rot", "temp", "pressure", "density",
                     "virial", "virial_bonded", "target_temp"):
            st[name] = getattr(s, name)
        st["dof"] = s.dof
        st.scal[SIDX["hb0"]:SIDX["hb2"] + 1] = list(s.hb)
        st.scal[SIDX["eta0"]:SIDX["eta2"] + 1] = list(s.eta)
        st.scal[SIDX["peta0"]:SIDX["peta2"] + 1] = list(s.p_eta)

    def call(self, fname, st: State, *args):
        s = self._sys(st)
        fn = getattr(self.lib, fname)
        r = fn(C.byref(s), *args)
        self._back(s, st)
        return r

    # -- RNG --
    def philox_words(self, ctr):
        c = (C.c_uint32 * 4)(*[int(x) & 0xFFFFFFFF for x in ctr])
        o = (C.c_uint32 * 4)()
        self.lib.or_philox4x32_10(c, o)
        return np.array(list(o), dtype=np.uint32)

    def philox_words_key(self, ctr, key):
        c = (C.c_uint32 * 4)(*[int(x) & 0xFFFFFFFF for x in ctr])
        k = (C.c_uint32 * 2)(*[int(x) & 0xFFFFFFFF for x in key])
        o = (C.c_uint32 * 4)()
        self.lib.or_philox4x32_10_key(c, k, o)
        return np.array(list(o), dtype=np.uint32)

    def philox_normal(self, c1, c2, c3, c4):
        o = [C.c_double() for _ in range(3)]
        self.lib.or_philox_normal_3seq(c1, c2, c3, c4, *[C.byref(x) for x in o])
        return np.array([x.value for x in o])

    # -- pair kernels --
    def pair(self, kind, st: State, i, j, p):
        i, j = int(i), int(j)
        s = self._sys(st)
        pe = C.c_double()
        f = np.zeros(3); ti = np.zeros(3); tj = np.zeros(3)
        d = C.c_double
        fp, tip, tjp = _dp(f), _dp(ti), _dp(tj)
        L = self.lib
        if kind == LJ:
            L.or_lj(C.byref(s), i, j, d(p[0]), d(p[1]), d(p[2]), C.byref(pe), fp)
        elif kind == LJ12:
            L.or_lj12(C.byref(s), i, j, d(p[0]), d(p[1]), d(p[2]), C.byref(pe), fp)
        elif kind == LJW:
            L.or_ljw(C.byref(s), i, j, d(p[0]), d(p[1]), d(p[2]), d(p[3]), C.byref(pe), fp)
        elif kind == WCA:
            L.or_wca(C.byref(s), i, j, d(p[0]), d(p[1]), C.byref(pe), fp)
        elif kind == MORSE:
            L.or_morse(C.byref(s), i, j, d(p[0]), d(p[1]), d(p[2]), d(p[3]), C.byref(pe), fp)
        elif kind == DH:
            L.or_dh(C.byref(s), i, j, d(p[0]), d(p[1]), d(p[2]), C.byref(pe), fp)
        elif kind == COULOMB:
            L.or_coulomb(C.byref(s), i, j, d(p[0]), d(p[1]), C.byref(pe), fp)
        elif kind in (GB, GBR):
            cut = p[0] if kind == GB else 0.0
            mu = int(p[1]) if len(p) > 1 else 1
            nu = int(p[2]) if len(p) > 2 else 2
            L.or_gb(C.byref(s), i, j, 0 if kind == GB else 1, d(cut), mu, nu,
                    C.byref(pe), fp, tip, tjp)
        elif kind in (GBW, GBC, GBWC):
            # p = cutoff, mu, nu, then wf (GBW, GBWC) and chirality (GBC, GBWC)
            wf = p[3] if kind in (GBW, GBWC) else 0.0
            ch = p[3] if kind == GBC else (p[4] if kind == GBWC else 0.0)
            L.or_gbx(C.byref(s), i, j, {GBW: 2, GBC: 3, GBWC: 4}[kind], d(p[0]), d(wf), d(ch), int(p[1]), int(p[2]),
                     C.byref(pe), fp, tip, tjp)
        elif kind in (ECP, ECPR, ECPW, ECPC, ECPWC):
            # p = cutoff, then wf (ECPW, ECPWC) and chirality (ECPC, ECPWC)
            cut = 0.0 if kind == ECPR else p[0]
            wf = p[1] if kind in (ECPW, ECPWC) else 0.0
            ch = p[1] if kind == ECPC else (p[2] if kind == ECPWC else 0.0)
            lam = C.c_double(); nit = C.c_int()
            L.or_ecpx(C.byref(s), i, j, {ECP: 0, ECPR: 1, ECPW: 2, ECPC: 3, ECPWC: 4}[kind], d(cut), d(wf), d(ch),
                      C.byref(pe), fp, tip, tjp, C.byref(lam), C.byref(nit))
            return pe.value, f, ti, tj, lam.value, nit.value
        else:
            raise ValueError(kind)
        return pe.value, f, ti, tj

    def ecp_shape(self, st: State, i, j):
        """(AR, BR, RIJ) of a pair: column-major 3x3 shape matrices and the
        min-image separation."""
        s = self._sys(st)
        ar = np.zeros(9); br = np.zeros(9); r = np.zeros(3)
        self.lib.or_ecp_shape(C.byref(s), int(i), int(j), _dp(ar), _dp(br), _dp(r))
        return ar, br, r

    def optim_dist(self, lam, ar, br, r):
        return self.lib.or_ecp_optim_dist(C.c_double(lam), _dp(ar), _dp(br), _dp(r))

    def lambda_converged(self, ar, br, r, l0=0.3819660112501051, tol=1e-10):
        """Converged contact parameter (safeguarded Newton) and its iterations."""
        it = C.c_int()
        f = self.lib.or_ecp_lambda_converged
        f.restype = C.c_double
        ar, br, r = (np.ascontiguousarray(x, dtype=np.float64) for x in (ar, br, r))
        lam = f(_dp(ar), _dp(br), _dp(r), C.c_double(l0), C.c_double(tol), C.byref(it))
        return lam, it.value

    def contact_distance(self, ar, br, r, lam):
        f = self.lib.or_ecp_contact_distance
        f.restype = C.c_double
        ar, br, r = (np.ascontiguousarray(x, dtype=np.float64) for x in (ar, br, r))
        return f(_dp(ar), _dp(br), _dp(r), C.c_double(lam))

    def features(self, st: State, i, j):
        """The 7 network inputs of pair (i, j)."""
        s = self._sys(st)
        x = np.zeros(7)
        self.lib.or_ecp_features(C.byref(s), int(i), int(j), _dp(x))
        return x

    def ecp_at_lambda(self, kind, st: State, i, j, p, lam):
        """ECP-family pair terms evaluated at a given contact parameter."""
        s = self._sys(st)
        pe = C.c_double()
        f = np.zeros(3); ti = np.zeros(3); tj = np.zeros(3)
        cut = 0.0 if kind == ECPR else p[0]
        wf = p[1] if kind in (ECPW, ECPWC) else 0.0
        ch = p[1] if kind == ECPC else (p[2] if kind == ECPWC else 0.0)
        d = C.c_double
        self.lib.or_ecpx_at_lambda(C.byref(s), int(i), int(j), {ECP: 0, ECPR: 1, ECPW: 2, ECPC: 3, ECPWC: 4}[kind],
                                   d(cut), d(wf), d(ch), d(lam), C.byref(pe), _dp(f), _dp(ti), _dp(tj))
        return pe.value, f, ti, tj

    def bonded(self, kind, st: State, idx, coeff, angle=0.0):
        idx = [int(x) for x in idx]
        s = self._sys(st)
        pe = C.c_double()
        f = np.zeros((4, 3))
        d = C.c_double
        rows = [_dp(f[k]) for k in range(4)]
        L = self.lib
        if kind == 1:
            L.or_bend(C.byref(s), idx[0], idx[1], idx[2], d(coeff), C.byref(pe), *rows[:3])
        elif kind == 2:
            L.or_cosine_bend(C.byref(s), idx[0], idx[1], idx[2], d(coeff), d(angle), C.byref(pe), *rows[:3])
        elif kind == 3:
            L.or_dih(C.byref(s), *idx[:4], d(coeff), C.byref(pe), *rows)
        elif kind == 4:
            L.or_cosine_dih(C.byref(s), *idx[:4], d(coeff), d(angle), C.byref(pe), *rows)
        elif kind == 5:
            L.or_angle_bend(C.byref(s), idx[0], idx[1], idx[2], d(coeff), d(angle), C.byref(pe), *rows[:3])
        elif kind == 6:
            L.or_angle_dih(C.byref(s), *idx[:4], d(coeff), d(angle), C.byref(pe), *rows)
        else:
            raise ValueError(kind)
        return pe.value, f

    def pair_loop(self, st: State, nl: dict, kind, p):
        s = self._sys(st)
        nls = self._nl(st, nl)
        pa = np.asarray(p, dtype=np.float64)
        self.lib.or_pair_loop(C.byref(s), C.byref(nls), kind, _dp(pa))
        self._back(s, st)

    # -- all-cores variants (liboracle_omp.so): the same results --
    def threads(self):
        return int(self.lib.or_threads())

    def pair_loop_threads(self, st: State, nl: dict, kind, p):
        s = self._sys(st)
        nls = self._nl(st, nl)
        pa = np.asarray(p, dtype=np.float64)
        buf = nl.get("_buf")
        if buf is None or len(buf) < 10 * max(nl["nnz"], 1):
            buf = nl["_buf"] = np.zeros(10 * max(nl["nnz"], 1) + 1024)
        self.lib.or_pair_loop_threads(C.byref(s), C.byref(nls), kind, _dp(pa), _dp(buf))
        self._back(s, st)

    def nlist_build_threads(self, st: State, nl: dict):
        cnt = np.zeros(st.n + 1, np.int64)
        while True:
            s = self._sys(st)
            o = self._nl(st, nl)
            r = self.lib.or_nlist_build_threads(C.byref(s), C.byref(o), cnt.ctypes.data_as(_PL))
            if r == 0:
                nl["nnz"] = o.nnz
                nl["r_skin_sq"] = o.r_skin_sq
                return nl
            nl["nbr"] = np.zeros(2 * len(nl["nbr"]), np.int32)

    # -- neighbour list --
    def _nl(self, st, nl):
        o = OrNlist()
        o.n = st.n
        o.cap = len(nl["nbr"])
        o.nnz = nl.get("nnz", 0)
        o.offset = nl["offset"].ctypes.data_as(_PL)
        o.nbr = nl["nbr"].ctypes.data_as(_PI)
        o.x_last = _dp(nl["x_last"]); o.y_last = _dp(nl["y_last"]); o.z_last = _dp(nl["z_last"])
        o.cutoff = nl["cutoff"]; o.skin = nl["skin"]; o.r_skin_sq = nl.get("r_skin_sq", 0.0)
        return o

    def nlist_new(self, st: State, cutoff, skin, cap=None):
        cap = cap or max(50 * st.n, 16)
        return {"offset": np.zeros(st.n + 1, np.int64), "nbr": np.zeros(cap, np.int32),
                "x_last": np.zeros(st.n), "y_last": np.zeros(st.n), "z_last": np.zeros(st.n),
                "cutoff": float(cutoff), "skin": float(skin), "nnz": 0}

    def nlist_build(self, st: State, nl: dict):
        while True:
            s = self._sys(st)
            o = self._nl(st, nl)
            r = self.lib.or_nlist_build(C.byref(s), C.byref(o))
            if r == 0:
                nl["nnz"] = o.nnz
                nl["r_skin_sq"] = o.r_skin_sq
                return nl
            nl["nbr"] = np.zeros(2 * len(nl["nbr"]), np.int32)

    def needs_rebuild(self, st: State, nl: dict):
        s = self._sys(st)
        o = self._nl(st, nl)
        return bool(self.lib.or_nlist_needs_rebuild(C.byref(s), C.byref(o)))

    # -- composite rigid bodies --
    def copy_transform(self, st: State, frm, to, offset, set_inertia=None):
        s = self._sys(st)
        d = C.c_double
        si = -1 if set_inertia is None else int(bool(set_inertia))
        r = self.lib.or_copy_transform(C.byref(s), int(frm), int(to), d(offset[0]), d(offset[1]), d(offset[2]), si)
        self._back(s, st)
        return r

    def transfer_forces(self, st: State, frm, to):
        s = self._sys(st)
        self.lib.or_transfer_forces(C.byref(s), int(frm), int(to))
        self._back(s, st)

    # -- point dipoles with a reaction field (module dipole) --
    def dipole_pair(self, st: State, i, j, cutoff, dielectric, dip, rfl):
        """dipole_calculate_forces for one pair; dip and rfl are (3, n)
        arrays updated in place (the module's DIP* and RFL*)."""
        s = self._sys(st)
        pe = C.c_double()
        f = np.zeros(3); ti = np.zeros(3); tj = np.zeros(3)
        self.lib.or_dipole(C.byref(s), int(i), int(j), C.c_double(cutoff), C.c_double(dielectric), C.byref(pe),
                           _dp(f), _dp(ti), _dp(tj), _dp(dip), _dp(rfl))
        return pe.value, f, ti, tj

    def dipole_loop(self, st: State, nl: dict, cutoff, dielectric, dip, rfl):
        s = self._sys(st)
        nls = self._nl(st, nl)
        self.lib.or_dipole_loop(C.byref(s), C.byref(nls), C.c_double(cutoff), C.c_double(dielectric), _dp(dip),
                                _dp(rfl))
        self._back(s, st)

    def eval_reaction_field(self, st: State, dip, rfl):
        s = self._sys(st)
        self.lib.or_eval_reaction_field(C.byref(s), _dp(dip), _dp(rfl))
        self._back(s, st)

    # -- Monte Carlo barostat --
    def mc_barostat(self, st: State, potential, target_pres, u, volume_change=None):
        """mc_barostat (integrators.f90:864-945) step by step, with the
        reference's expressions in its order; potential(st) is its
        compute_potential (it sets st's forces and returns pot_eng); u are the
        two numbers the move draws with random_number.  Returns accept
        (None where the reference returns early without setting it)."""
        if target_pres <= 1e-3:
            raise ValueError("Error: Target pressure too small.")
        max_dV = 0.01 if volume_change is None else volume_change
        beta = 1.0 / st["temp"]
        V_old = st.box[0] * st.box[1] * st.box[2]
        L_old = st.box.copy()
        for a, b in (("rx_old", "rx"), ("ry_old", "ry"), ("rz_old", "rz")):
            st[a] = st[b].copy()
        U_old = potential(st)
        V_new = V_old + max_dV * (2.0 * u[0] - 1.0) * V_old
        if V_new <= 0.0:
            return None
        scale = (V_new / V_old) ** (1.0 / 3.0)
        st.box[:] = L_old * scale
        for a, b in (("rx", "rx_old"), ("ry", "ry_old"), ("rz", "rz_old")):
            st[a] = st[b] * scale
        U_new = potential(st)
        dU = U_new - U_old
        exponent = -beta * dU - beta * target_pres * (V_new - V_old) + float(st.n) * np.log(V_new / V_old)
        if np.log(u[1]) < exponent:
            volume = st.box[0] * st.box[1] * st.box[2]
            st["density"] = float(st.n) / volume
            return True
        st.box[:] = L_old
        for a, b in (("rx", "rx_old"), ("ry", "ry_old"), ("rz", "rz_old")):
            st[a] = st[b].copy()
        return False

    # -- integrators --
    def integrate(self, kind, phase, st: State, dt, step=0, damping=1.0, coupling=None):
        d = C.c_double
        cp = C.byref(C.c_double(coupling)) if coupling is not None else None
        names = {VV: ("or_vv_initial_step", "or_vv_final_step"),
                 VVC: ("or_vvc_initial_step", "or_vvc_final_step"),
                 QQ: ("or_qq_initial_step", "or_qq_final_step"),
                 NHT: ("or_nht_initial_step", "or_nht_final_step"),
                 NHTC: ("or_nhtc_initial_step", "or_nhtc_final_step"),
                 LGV: ("or_lgv_initial_step", "or_lgv_final_step"),
                 LGVC: ("or_lgvc_initial_step", "or_lgvc_final_step")}
        name = names[kind][phase]
        if kind in (NHT, NHTC):
            return self.call(name, st, d(dt), cp)
        if kind in (LGV, LGVC):
            if phase == 0:
                return self.call(name, st, d(dt), C.c_int(step), d(damping))
            return self.call(name, st, d(dt), d(damping))
        return self.call(name, st, d(dt))


# ---------------------------------------------------------------------------
# the reference, built from source (oracle/_ref)
# ---------------------------------------------------------------------------
def reference_path(fast=False):
    return os.path.join(HERE, "_ref", "libsucrose_ref_fast.so" if fast else "libsucrose_ref.so")


def reference_available(fast=False):
    return os.path.exists(reference_path(fast))


class Reference:
    """The SUCROSE Fortran modules themselves, through oracle/ref_harness.f90.

    The library holds one system in module globals, so each instance drives
    that single global system."""

    def __init__(self, fast=False, private=False):
        path = reference_path(fast)
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} not built (make -C oracle ref)")
        if private:
            # a private copy of the library gets its own module globals (the
            # reference's neighbour list can be initialised once per copy)
            import shutil
            import tempfile
            fd, copy = tempfile.mkstemp(suffix=".so")
            os.close(fd)
            shutil.copyfile(path, copy)
            path = copy
        self.lib = C.CDLL(path)
        L = self.lib
        L.ref_brent.restype = C.c_double
        L.ref_brent.argtypes = [_PD, _PD, _PD]
        L.ref_needs_rebuild.restype = C.c_int
        L.ref_integrate.argtypes = [C.c_int, C.c_int, C.c_double, C.c_int, C.c_double, C.c_int, C.c_double]
        L.ref_run_steps.argtypes = [C.c_int, C.c_int, _PD, C.c_int, C.c_int, C.c_double, C.c_double]
        self.n = None

    def load(self, st: State):
        L = self.lib
        L.ref_alloc(C.c_int(st.n), C.c_int(st.nrigid))
        self.n = st.n
        L.ref_set_box(_dp(st.box), C.c_int(1 if st.periodic else 0))
        buf = np.asfortranarray(st.buf)
        L.ref_set_state(C.c_int(st.n), _dp(buf))
        sc = st.scal.copy()
        L.ref_set_scalars(_dp(sc))
        if st.nc > 0:
            bi = np.ascontiguousarray(st.bbi + 1, dtype=np.int32)
            bj = np.ascontiguousarray(st.bbj + 1, dtype=np.int32)
            bl = np.ascontiguousarray(st.bb_len, dtype=np.float64)
            L.ref_set_bonds(C.c_int(st.nc), bi.ctypes.data_as(_PI), bj.ctypes.data_as(_PI), _dp(bl))

    def store(self, st: State):
        buf = np.zeros((st.n, len(FIELDS)), order="F")
        self.lib.ref_get_state(C.c_int(st.n), _dp(buf))
        st.buf[:] = buf
        sc = np.zeros(len(SCALARS))
        self.lib.ref_get_scalars(_dp(sc))
        st.scal[:] = sc

    def philox_uniform(self, ctr):
        c = (C.c_int32 * 4)(*[int(np.int32(np.uint32(x))) for x in ctr])
        o = np.zeros(4)
        self.lib.ref_philox_uniform(c, _dp(o))
        return o

    def philox_normal(self, c1, c2, c3, c4):
        o = np.zeros(3)
        self.lib.ref_philox_normal(C.c_int32(c1), C.c_int32(c2), C.c_int32(c3), C.c_int32(c4), _dp(o))
        return o

    def cholesky_solve(self, a, b):
        a = np.asfortranarray(a, dtype=np.float64); b = np.ascontiguousarray(b, dtype=np.float64)
        o = np.zeros(3)
        self.lib.ref_cholesky_solve(_dp(a), _dp(b), _dp(o))
        return o

    def brent(self, ar, br, r):
        ar = np.asfortranarray(ar, dtype=np.float64); br = np.asfortranarray(br, dtype=np.float64)
        r = np.ascontiguousarray(r, dtype=np.float64)
        return self.lib.ref_brent(_dp(ar), _dp(br), _dp(r))

    def pair(self, kind, i, j, p):
        i, j = int(i), int(j)
        pa = np.zeros(8)
        pa[:len(p)] = p
        pe = C.c_double()
        f = np.zeros(3); ti = np.zeros(3); tj = np.zeros(3)
        self.lib.ref_pair(C.c_int(kind), C.c_int(i + 1), C.c_int(j + 1), _dp(pa), C.byref(pe),
                          _dp(f), _dp(ti), _dp(tj))
        return pe.value, f, ti, tj

    def report(self, kind, path, step, rtype):
        """One frame of reporters.f90 written by the reference to a new file
        (kind 1 state data, 2 trajectory, 3 config, 4 checkpoint)."""
        p = os.fsencode(path)
        rt = np.ascontiguousarray(rtype, dtype=np.int32)
        self.lib.ref_report(C.c_int(kind), p, C.c_int(len(p)), C.c_int(step), rt.ctypes.data_as(_PI))

    def load_checkpoint(self, path):
        """load_checkpoint (reporters.f90:188-221) into the module arrays;
        returns n, nc, config_index, the field buffer, rtype, bonds, box."""
        p = os.fsencode(path)
        n, nc, cfg = C.c_int(), C.c_int(), C.c_int()
        self.lib.ref_load_checkpoint(p, C.c_int(len(p)), C.byref(n), C.byref(nc), C.byref(cfg))
        self.n = n.value
        buf = np.zeros((n.value, len(FIELDS)), order="F")
        self.lib.ref_get_state(C.c_int(n.value), _dp(buf))
        rt = np.zeros(n.value, np.int32)
        self.lib.ref_get_types(rt.ctypes.data_as(_PI))
        m = max(nc.value, 1)
        bi = np.zeros(m, np.int32); bj = np.zeros(m, np.int32); bl = np.zeros(m)
        self.lib.ref_get_bonds(bi.ctypes.data_as(_PI), bj.ctypes.data_as(_PI), _dp(bl))
        box = np.zeros(3)
        self.lib.ref_get_box(_dp(box))
        return n.value, nc.value, cfg.value, np.ascontiguousarray(buf), rt, (bi[:nc.value], bj[:nc.value],
                                                                             bl[:nc.value]), box

    def copy_transform(self, frm, to, offset, set_inertia=None):
        d = C.c_double
        si = -1 if set_inertia is None else int(bool(set_inertia))
        self.lib.ref_copy_transform(C.c_int(frm + 1), C.c_int(to + 1), d(offset[0]), d(offset[1]), d(offset[2]),
                                    C.c_int(si))

    def transfer_forces(self, frm, to):
        self.lib.ref_transfer_forces(C.c_int(frm + 1), C.c_int(to + 1))

    def init_reaction_field(self):
        self.lib.ref_init_reaction_field()

    def dipole_pair(self, i, j, cutoff, dielectric):
        pe = C.c_double()
        f = np.zeros(3); ti = np.zeros(3); tj = np.zeros(3)
        self.lib.ref_dipole_pair(C.c_int(i + 1), C.c_int(j + 1), C.c_double(cutoff), C.c_double(dielectric),
                                 C.byref(pe), _dp(f), _dp(ti), _dp(tj))
        return pe.value, f, ti, tj

    def eval_reaction_field(self):
        self.lib.ref_eval_reaction_field()

    def get_dipole(self):
        dip = np.zeros((3, self.n)); rfl = np.zeros((3, self.n))
        d = np.zeros((self.n, 3), order="F"); r = np.zeros((self.n, 3), order="F")
        self.lib.ref_get_dipole(_dp(d), _dp(r))
        return np.ascontiguousarray(d.T), np.ascontiguousarray(r.T)

    def mc_barostat(self, seed, target_pres, volume_change, sigma, eps, rc):
        """mc_barostat with the harness's all-pairs Lennard-Jones potential;
        returns the two random numbers it drew and accept."""
        u = np.zeros(2)
        acc = C.c_int()
        d = C.c_double
        self.lib.ref_mc_barostat(C.c_int(seed), d(target_pres), d(volume_change), d(sigma), d(eps), d(rc), _dp(u),
                                 C.byref(acc))
        return u, bool(acc.value)

    def bonded(self, kind, idx, coeff, angle=0.0):
        idx = list(idx) + [0] * (4 - len(idx))
        pe = C.c_double()
        f = np.zeros((4, 3))
        self.lib.ref_bonded(C.c_int(kind), *[C.c_int(x + 1) for x in idx], C.c_double(coeff),
                            C.c_double(angle), C.byref(pe), _dp(f))
        return pe.value, f

    def integrate(self, kind, phase, dt, step=0, damping=1.0, coupling=None):
        self.lib.ref_integrate(kind, phase, dt, step, damping,
                               0 if coupling is None else 1, 0.0 if coupling is None else coupling)

    def constraints(self, phase, dt):
        self.lib.ref_constraints(C.c_int(phase), C.c_double(dt))

    def initialize_system(self, remove_com=True):
        self.lib.ref_initialize_system(C.c_int(1 if remove_com else 0))

    def initialize_forces(self):
        self.lib.ref_initialize_forces()

    def calculate_thermo(self):
        self.lib.ref_calculate_thermo()

    def set_velocity_to_temperature(self, t, seed, remove_com=True):
        self.lib.ref_set_velocity_to_temperature(C.c_double(t), C.c_int(seed), C.c_int(1 if remove_com else 0))

    def generate_fcc_lattice(self, uc, dens):
        self.lib.ref_generate_fcc_lattice(C.c_int(uc), C.c_double(dens))

    def nlist_init(self, skin, cutoff):
        self.lib.ref_nlist_init(C.c_double(skin), C.c_double(cutoff))

    def nlist_build(self):
        self.lib.ref_nlist_build()

    def needs_rebuild(self):
        return bool(self.lib.ref_needs_rebuild())

    def nlist_get(self, cap=None):
        cap = cap or 50 * self.n * 4
        off = np.zeros(self.n, np.int32)
        nbr = np.zeros(cap, np.int32)
        nnz = C.c_long()
        self.lib.ref_nlist_get(C.c_long(cap), off.ctypes.data_as(_PI), nbr.ctypes.data_as(_PI), C.byref(nnz))
        return off, nbr[:nnz.value].copy(), nnz.value

    def run_steps(self, integ, fkind, p, nsteps, step0, dt, damping=1.0):
        pa = np.zeros(8)
        pa[:len(p)] = p
        self.lib.ref_run_steps(integ, fkind, _dp(pa), nsteps, step0, dt, damping)

    def run_chain_steps(self, nsteps, step0, dt, damping, wca, dh, bends0, kbend, abend, dihs0, kdih):
        """The charged-chain loop of ref_harness.f90 (lgvc + WCA and DH over
        the list + cosine_bend_ and dih_ terms); bends0 (m, 3) and dihs0 (m, 4)
        hold 0-based atoms."""
        pw = np.asarray(wca, dtype=np.float64)
        pd = np.asarray(dh, dtype=np.float64)
        b = np.ascontiguousarray(np.asarray(bends0) + 1, dtype=np.int32)
        d = np.ascontiguousarray(np.asarray(dihs0) + 1, dtype=np.int32)
        self.lib.ref_run_chain_steps(C.c_int(nsteps), C.c_int(step0), C.c_double(dt), C.c_double(damping),
                                     _dp(pw), _dp(pd), C.c_int(len(b)), b.ctypes.data_as(_PI),
                                     C.c_double(kbend), C.c_double(abend), C.c_int(len(d)),
                                     d.ctypes.data_as(_PI), C.c_double(kdih))
