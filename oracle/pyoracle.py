"""ctypes wrapper of the FP64 CPU oracle (mpm_oracle.c) -- TEST INFRASTRUCTURE.

Only tests, the smoke check and the CPU-baseline leg of the benchmark use
this module, as the checker. It also wraps oracle/_ref/libmpm_ref.so, the
harness around the reference's own constitutive_model.cpp, bc.cpp and
cal_mat.h / map_and_interpolate.h templates, when that library was built.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_build", "libmpm_oracle.so")
LIB_FAST = os.path.join(HERE, "_build", "libmpm_oracle_fast.so")
LIB_OMP = os.path.join(HERE, "_build", "libmpm_oracle_omp.so")  # multithreaded CPU-baseline port
REF_LIB = os.path.join(HERE, "_ref", "libmpm_ref.so")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)

PIC, FLIP, TPIC, APIC = 0, 1, 2, 3


class OrcState(C.Structure):
    _fields_ = [
        ("xmin", C.c_double * 3), ("xmax", C.c_double * 3), ("dx", C.c_double * 3),
        ("nex", C.c_int32 * 3), ("idimc", C.c_int32 * 3), ("nxc", C.c_int32 * 3),
        ("nelem", C.c_int32), ("nodec", C.c_int32),
        ("dt", C.c_double), ("mtol", C.c_double), ("bb", C.c_double * 3),
        ("scheme", C.c_int32), ("nlstep", C.c_int32), ("shifting", C.c_int32), ("pad0", C.c_int32),
        ("nmat", C.c_int32), ("iprop", C.c_int32 * 8), ("mat_prop", (C.c_double * 8) * 8),
        ("ibc", C.c_int32 * 4), ("nbc", _ip * 4), ("fbc", _dp * 4),
        ("num", C.c_int32),
        ("coord", _dp), ("vel", _dp), ("mass", _dp), ("vol", _dp), ("vol0", _dp), ("stress", _dp),
        ("def_grad", _dp), ("det_def_grad", _dp), ("trac_force", _dp), ("matid", _ip), ("eps_p", _dp),
        ("idepf", _ip), ("idp2p", _ip), ("numep", _ip),
        ("nmass", _dp), ("nvof", _dp), ("nmome", _dp), ("nforce", _dp), ("nvel", _dp), ("naccel", _dp),
        ("nvol", _dp), ("ddg", _dp), ("cmat", _dp), ("dinv", _dp),
        ("fbar", C.c_int32), ("pad1", C.c_int32), ("def_grad_bar", _dp), ("det_def_grad_bar", _dp),
        ("njbar", _dp), ("det_ddg", _dp),
    ]


def build():
    subprocess.run(["make", "-s", "-C", HERE, "oracle", "timing", "omp"], check=True)


def mutant_path(k):
    """Parity build with planted mistake k in orc_shift_delta (1: no factor 2
    in the gradient, 2: vol0 for vol), for tests that must reject it."""
    return os.path.join(HERE, "_build", "libmpm_oracle_mut%d.so" % k)


def load(fast=False):
    """fast=False: parity build; True: timing build; "omp": timing build with
    the OpenMP port of the step (orc_run_omp); ("mutant", k): mutant_path(k)."""
    if isinstance(fast, tuple) and fast[0] == "mutant":
        path = mutant_path(fast[1])
        if not os.path.exists(path):
            subprocess.run(["make", "-s", "-C", HERE, "mutants"], check=True)
    else:
        path = LIB_OMP if fast == "omp" else (LIB_FAST if fast else LIB)
    if not os.path.exists(path):
        build()
    lib = C.CDLL(path)
    lib.orc_grid_init.argtypes = [C.POINTER(OrcState)]
    lib.orc_maksf.argtypes = [C.POINTER(OrcState), C.c_int, _dp, _ip, _dp, _dp]
    lib.orc_maksf_batch.argtypes = [C.POINTER(OrcState), C.c_int, _ip, _dp, _ip, _dp, _dp]
    lib.orc_cell_of.argtypes = [C.POINTER(OrcState), _dp]
    for f in ("orc_bin", "orc_solve_solid"):
        getattr(lib, f).argtypes = [C.POINTER(OrcState)]
    lib.orc_p2g.argtypes = [C.POINTER(OrcState), C.c_double]
    lib.orc_g2p.argtypes = [C.POINTER(OrcState), C.c_double]
    lib.orc_step.argtypes = [C.POINTER(OrcState), C.c_int]
    lib.orc_run.argtypes = [C.POINTER(OrcState), C.c_int, C.c_int]
    lib.orc_facl.argtypes = [C.POINTER(OrcState), C.c_int]
    lib.orc_facl.restype = C.c_double
    lib.orc_update_stress.argtypes = [C.c_int, _dp, _dp, _dp, _dp, C.c_double, C.c_double]
    lib.orc_j2_return.argtypes = [_dp, _dp, _dp, _dp, _dp, _dp]
    lib.orc_det3.argtypes = [_dp]
    lib.orc_det3.restype = C.c_double
    lib.orc_inv3.argtypes = [_dp, _dp]
    lib.orc_bc_set_val.argtypes = [C.POINTER(OrcState), C.c_int, C.c_int, _dp, C.c_double]
    lib.orc_bc_set_zero.argtypes = [C.POINTER(OrcState), C.c_int, C.c_int, _dp]
    lib.orc_tait_stress.argtypes = [_dp, C.c_double, _dp, C.c_double, _dp]
    lib.orc_nodal_vol.argtypes = [C.POINTER(OrcState), C.POINTER(C.c_int32)]
    lib.orc_shift_delta.argtypes = [C.POINTER(OrcState), _dp]
    lib.orc_node_coord.argtypes = [C.POINTER(OrcState), C.c_int, C.c_int]
    lib.orc_node_coord.restype = C.c_double
    lib.orc_fbar.argtypes = [C.POINTER(OrcState)]
    lib.orc_newmark_para.argtypes = [C.c_double, C.c_double, C.c_double, _dp]
    lib.orc_implicit_p2g.argtypes = [C.POINTER(OrcState), _dp, C.c_double]
    lib.orc_implicit_g2p.argtypes = [C.POINTER(OrcState), _dp, _dp, _dp, C.c_double]
    if fast == "omp":
        lib.orc_run_omp.argtypes = [C.POINTER(OrcState), C.c_int, C.c_int, C.c_int]
    _fp = C.POINTER(C.c_float)
    lib.orc_mlp_forward.argtypes = [_dp, _dp, _fp, C.c_long, C.c_int, _dp]
    lib.orc_mlp_grad.argtypes = [_dp, _dp, _fp, _fp, C.c_long, C.c_int, _dp]
    lib.orc_mlp_grad.restype = C.c_double
    return lib


MLP_NPARAMS = 6288


def newmark_para(gamma_nb, beta_nb, dt):
    """The six Newmark coefficients of NewmarkBetaParaSet."""
    nb = np.zeros(6)
    load().orc_newmark_para(gamma_nb, beta_nb, dt, P(nb))
    return nb


def mlp_forward(params, scaling, inp, bf16=True):
    """Oracle forward pass: params (6288,), scaling (64,), inp (n, 16) float32 -> (n, 16) float64."""
    lib = load()
    p = np.ascontiguousarray(params, dtype=np.float64)
    sc = np.ascontiguousarray(scaling, dtype=np.float64)
    x = np.ascontiguousarray(inp, dtype=np.float32)
    out = np.zeros((x.shape[0], 16), dtype=np.float64)
    lib.orc_mlp_forward(P(p), P(sc), P(x, C.c_float), x.shape[0], int(bf16), P(out))
    return out


def mlp_grad(params, scaling, inp, tgt, bf16=True):
    """Oracle loss and gradient (6288,) of the mean squared error in normalised units."""
    lib = load()
    p = np.ascontiguousarray(params, dtype=np.float64)
    sc = np.ascontiguousarray(scaling, dtype=np.float64)
    x = np.ascontiguousarray(inp, dtype=np.float32)
    t = np.ascontiguousarray(tgt, dtype=np.float32)
    g = np.zeros(MLP_NPARAMS, dtype=np.float64)
    loss = lib.orc_mlp_grad(P(p), P(sc), P(x, C.c_float), P(t, C.c_float), x.shape[0], int(bf16), P(g))
    return loss, g


def load_ref():
    """The reference-numerics harness, or None when it was not built."""
    if not os.path.exists(REF_LIB):
        return None
    lib = C.CDLL(REF_LIB)
    lib.ref_update_stress.argtypes = [C.c_int, _dp, _dp, _dp, _dp, C.c_int, C.c_double, C.c_double]
    lib.ref_bc_apply.argtypes = [C.c_int, C.c_int, _ip, _dp, C.c_int, _dp, C.c_int, C.c_double, C.c_double]
    lib.ref_det3.argtypes = [_dp]
    lib.ref_det3.restype = C.c_double
    lib.ref_inv3.argtypes = [_dp, _dp]
    lib.ref_var_g2p.argtypes = [C.c_int, C.c_int, _dp, C.c_int, _dp, C.c_int, _ip, _ip, _dp, C.c_double]
    lib.ref_trace3.argtypes = [_dp, _dp]
    return lib


def P(a, ct=C.c_double):
    return a.ctypes.data_as(C.POINTER(ct))


class Oracle:
    """Holds an orc_state and the numpy arrays it points into."""

    def __init__(self, scene, fast=False):
        self.lib = load(fast)
        st = OrcState()
        for i in range(3):
            st.xmin[i] = scene.xmin[i]
            st.xmax[i] = scene.xmax[i]
            st.nex[i] = scene.nelem[i]
            st.idimc[i] = int(getattr(scene, "idimc", (1, 1, 1))[i])
        self.lib.orc_grid_init(C.byref(st))
        st.dt = scene.dt
        st.mtol = scene.mtol
        for i in range(3):
            st.bb[i] = scene.bb[i]
        st.scheme = scene.scheme
        st.nlstep = scene.nlstep
        st.shifting = 0
        st.nmat = len(scene.materials)
        for i, (model, props) in enumerate(scene.materials):
            st.iprop[i] = model
            for k, v in enumerate(props):
                st.mat_prop[i][k] = v
        self._keep = []
        for comp in range(4):
            if comp in scene.bcs:
                nbc, fbc = scene.bcs[comp]
                nbc = np.ascontiguousarray(nbc, np.int32)
                fbc = np.ascontiguousarray(fbc, np.float64)
                self._keep += [nbc, fbc]
                st.ibc[comp] = len(nbc)
                st.nbc[comp] = P(nbc, C.c_int32)
                st.fbc[comp] = P(fbc)
            else:
                st.ibc[comp] = 0
        p = scene.parts.copy()
        self.parts = p
        n = p.n
        st.num = n
        for f in ("coord", "vel", "mass", "vol", "vol0", "stress", "def_grad", "det_def_grad", "eps_p"):
            setattr(st, f, P(getattr(p, f)))
        if p.trac_force is not None:
            st.trac_force = P(p.trac_force)
        if p.matid is not None:
            st.matid = P(p.matid, C.c_int32)
        nelem, nodec = st.nelem, st.nodec
        self.idepf = np.zeros(nelem, np.int32)
        self.idp2p = np.zeros(n, np.int32)
        self.numep = np.zeros(nelem, np.int32)
        st.idepf, st.idp2p, st.numep = P(self.idepf, C.c_int32), P(self.idp2p, C.c_int32), P(self.numep, C.c_int32)
        self.nmass = np.zeros(nodec)
        self.nvof = np.zeros(nodec)
        self.nmome = np.zeros(3 * nodec)
        self.nforce = np.zeros(3 * nodec)
        self.nvel = np.zeros(3 * nodec)
        self.naccel = np.zeros(3 * nodec)
        self.ddg = np.zeros(9 * n)
        for f in ("nmass", "nvof", "nmome", "nforce", "nvel", "naccel", "ddg"):
            setattr(st, f, P(getattr(self, f)))
        self.nvol = np.zeros(nodec)
        st.nvol = P(self.nvol)
        # TPIC velocity gradient / APIC B matrix and APIC inverse D, [n][3][3], zero at start
        self.cmat = np.zeros((n, 3, 3))
        self.dinv = np.zeros((n, 3, 3))
        st.cmat, st.dinv = P(self.cmat), P(self.dinv)
        # F-bar state (identity / one at start) and scratch
        self.def_grad_bar = np.tile(np.eye(3), (n, 1, 1))
        self.det_def_grad_bar = np.ones(n)
        self.njbar = np.zeros(nodec)
        self.det_ddg = np.zeros(n)
        for f in ("def_grad_bar", "det_def_grad_bar", "njbar", "det_ddg"):
            setattr(st, f, P(getattr(self, f)))
        st.fbar = 1 if getattr(scene, "fbar", False) else 0
        self.st = st
        self.nodec = nodec
        self.nelem = nelem
        if getattr(scene, "shifting", False):
            self.set_shifting(True, scene.npxye)

    def set_shifting(self, on, npxye=(2, 2, 2)):
        if on:
            npx = (C.c_int32 * 3)(*npxye)
            self.lib.orc_nodal_vol(C.byref(self.st), npx)
        self.st.shifting = 1 if on else 0

    def shift_delta(self):
        d = np.zeros((self.parts.n, 3))
        self.lib.orc_shift_delta(C.byref(self.st), P(d))
        return d

    def node_coord(self, axis, idx):
        return self.lib.orc_node_coord(C.byref(self.st), axis, idx)

    def fbar(self):
        self.lib.orc_fbar(C.byref(self.st))

    def bin(self):
        return self.lib.orc_bin(C.byref(self.st))

    def p2g(self, facl=1.0):
        self.lib.orc_p2g(C.byref(self.st), facl)

    def solve_grid(self):
        self.lib.orc_solve_solid(C.byref(self.st))

    def g2p(self, facl=1.0):
        self.lib.orc_g2p(C.byref(self.st), facl)

    def implicit_p2g(self, accel, facl=1.0):
        """Particle2Node of the implicit path; accel: (n, 3) particle accelerations."""
        accel = np.ascontiguousarray(accel, np.float64)
        self.lib.orc_implicit_p2g(C.byref(self.st), P(accel), facl)

    def implicit_g2p(self, accel, ndispl, nb_para, gamma_nb):
        """Node2Particle of the implicit path for nodal displacement ndispl
        ([u|v|w] blocks); accel ((n, 3), C-contiguous FP64) is updated in place."""
        assert accel.flags.c_contiguous and accel.dtype == np.float64
        nd = np.ascontiguousarray(ndispl, np.float64)
        nb = np.ascontiguousarray(nb_para, np.float64)
        return self.lib.orc_implicit_g2p(C.byref(self.st), P(accel), P(nd), P(nb), gamma_nb)

    def step(self, istep):
        return self.lib.orc_step(C.byref(self.st), istep)

    def run(self, ista, iend):
        return self.lib.orc_run(C.byref(self.st), ista, iend)

    def run_omp(self, ista, iend, threads=0):
        """The OpenMP port (needs Oracle(scene, fast="omp"))."""
        return self.lib.orc_run_omp(C.byref(self.st), ista, iend, threads)

    def cell_of(self, x):
        x = np.ascontiguousarray(x, np.float64)
        return self.lib.orc_cell_of(C.byref(self.st), P(x))
