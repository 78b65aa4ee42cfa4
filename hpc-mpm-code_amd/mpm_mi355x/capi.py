"""ctypes binding of the engine's C ABI (include/mpm/mpm_capi.h).

The structures below mirror the header field for field. The loaded library
is the in-tree ``libmpm_mi355x.so``; there is no fallback: if it is missing
or the GPU is absent, calls fail with an exception.
"""
import ctypes as C
import os

import numpy as np

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.environ.get("MPM_MI355X_LIB") or os.path.join(PKG_DIR, "libmpm_mi355x.so")
HEADER_PATH = os.path.join(os.path.dirname(PKG_DIR), "include", "mpm", "mpm_capi.h")

MPM_OK = 0
MPM_ERR_INVALID_ARGUMENT = -1
MPM_ERR_OUT_OF_GRID = -2
MPM_ERR_NONFINITE = -3
MPM_ERR_JACOBIAN = -4
MPM_ERR_HIP = -5
MPM_ERR_UNSUPPORTED = -6
MPM_ERR_NO_DEVICE = -7
MPM_ERR_INTERNAL = -8
MPM_ERR_HOST_MEMORY = -9
NOALLOC_MAX_POSITIONS = 1048576  # MPM_NOALLOC_MAX_POSITIONS

SCHEME_PIC, SCHEME_FLIP, SCHEME_TPIC, SCHEME_APIC = 0, 1, 2, 3
PREC_FP64_EXACT, PREC_FP32 = 0, 1
MODEL_RIGID, MODEL_LINELAST, MODEL_NEOHOOKEAN, MODEL_STVK, MODEL_MOONEYRIVLIN, MODEL_J2 = -1, 0, 1, 2, 3, 10
MODEL_TAIT = 20
OPT_FULL_NODE_OUTPUT = 1
OPT_RELAYOUT_EVERY = 3
OPT_GRAPH_BLOCK = 4
OPT_FUSED_COUNT = 5
OPT_LANES_PER_CELL = 2
OPT_CELLS_PREFETCH = 6
OPT_CELLS_GRID = 7
OPT_X_MERGE = 8
OPT_SURROGATE_FUSED = 9

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)


class GridDesc(C.Structure):
    _fields_ = [("xmin", C.c_double * 3), ("xmax", C.c_double * 3), ("nelem", C.c_int32 * 3),
                ("idimc", C.c_int32 * 3)]


class MaterialDesc(C.Structure):
    _fields_ = [("model", C.c_int32), ("nprop", C.c_int32), ("prop", C.c_double * 8)]


class StepParams(C.Structure):
    _fields_ = [("dt", C.c_double), ("mtol", C.c_double), ("bb", C.c_double * 3), ("scheme", C.c_int32),
                ("nlstep", C.c_int32), ("precision", C.c_int32), ("relayout_every", C.c_int32)]


class ParticlesView(C.Structure):
    _fields_ = [("n", C.c_int64), ("coord", _dp), ("vel", _dp), ("mass", _dp), ("vol", _dp), ("vol0", _dp),
                ("stress", _dp), ("def_grad", _dp), ("det_def_grad", _dp), ("trac_force", _dp), ("matid", _ip),
                ("eps_p", _dp)]


class NodesView(C.Structure):
    _fields_ = [("nodec", C.c_int64), ("nmass", _dp), ("nvof", _dp), ("nmome", _dp), ("nforce", _dp),
                ("nvel", _dp), ("naccel", _dp)]


class KernelTimes(C.Structure):
    _fields_ = [("bin_ms", C.c_double), ("p2g_ms", C.c_double), ("p2g_cells_ms", C.c_double),
                ("p2g_nodes_ms", C.c_double), ("g2p_vel_ms", C.c_double), ("musl_cells_ms", C.c_double),
                ("musl_nodes_ms", C.c_double), ("g2p_upd_ms", C.c_double), ("total_ms", C.c_double),
                ("steps", C.c_int32), ("active_cells", C.c_int32), ("active_nodes", C.c_int32),
                ("pad", C.c_int32)]


# name -> (restype, argtypes); every function declared in mpm_capi.h
class TrainParams(C.Structure):
    _fields_ = [("epochs", C.c_int32), ("batch", C.c_int32), ("lr", C.c_float), ("lr_final", C.c_float),
                ("seed", C.c_uint32), ("pad", C.c_int32)]


class SurrogateStats(C.Structure):
    _fields_ = [("samples", C.c_int64), ("logged_steps", C.c_int64), ("evaluated", C.c_int64),
                ("fallbacks", C.c_int64), ("loss", C.c_float), ("tol", C.c_float), ("offered", C.c_int64)]


class TuneResult(C.Structure):
    _fields_ = [("bucket", C.c_int32), ("lanes_per_cell", C.c_int32), ("relayout_every", C.c_int32),
                ("scenes", C.c_int32), ("ms_per_step", C.c_double), ("ms_per_step_worst", C.c_double)]


class EngineInfo(C.Structure):
    _fields_ = [("n", C.c_int64), ("nelem", C.c_int32), ("nodec", C.c_int32), ("precision", C.c_int32),
                ("scheme", C.c_int32), ("shifting", C.c_int32), ("fbar", C.c_int32)]


class CaseDesc(C.Structure):
    _fields_ = [("grid", GridDesc), ("step", StepParams), ("ista", C.c_int32), ("iend", C.c_int32),
                ("iout", C.c_int32), ("rstflag", C.c_int32), ("npxye", C.c_int32 * 3), ("nmat", C.c_int32),
                ("rho", C.c_double), ("n", C.c_int64), ("ibc", C.c_int32 * 3), ("pad", C.c_int32)]


_fp = C.POINTER(C.c_float)

SIGNATURES = {
    "mpm_version": (C.c_char_p, []),
    "mpm_last_error": (C.c_char_p, []),
    "mpm_engine_create": (C.c_int, [C.POINTER(GridDesc), C.POINTER(MaterialDesc), C.c_int32,
                                    C.POINTER(StepParams), C.POINTER(C.c_void_p)]),
    "mpm_engine_destroy": (C.c_int, [C.c_void_p]),
    "mpm_set_bc": (C.c_int, [C.c_void_p, C.c_int32, _ip, _dp, C.c_int32]),
    "mpm_upload_particles": (C.c_int, [C.c_void_p, C.POINTER(ParticlesView)]),
    "mpm_download_particles": (C.c_int, [C.c_void_p, C.POINTER(ParticlesView)]),
    "mpm_download_nodes": (C.c_int, [C.c_void_p, C.POINTER(NodesView)]),
    "mpm_bin": (C.c_int, [C.c_void_p]),
    "mpm_p2g": (C.c_int, [C.c_void_p, C.c_double]),
    "mpm_solve_grid": (C.c_int, [C.c_void_p]),
    "mpm_g2p": (C.c_int, [C.c_void_p, C.c_double]),
    "mpm_download_bins": (C.c_int, [C.c_void_p, _ip, _ip, _ip, _ip]),
    "mpm_step": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32]),
    "mpm_sync": (C.c_int, [C.c_void_p]),
    "mpm_get_kernel_times": (C.c_int, [C.c_void_p, C.POINTER(KernelTimes)]),
    "mpm_set_option": (C.c_int, [C.c_void_p, C.c_int32, C.c_int64]),
    "mpm_set_timing": (C.c_int, [C.c_void_p, C.c_int32]),
    "mpm_maksf_batch": (C.c_int, [C.POINTER(GridDesc), C.c_int64, _ip, _dp, _ip, _dp, _dp]),
    "mpm_bin_positions": (C.c_int, [C.POINTER(GridDesc), C.c_int64, _dp, _ip, _ip]),
    "mpm_set_material_dt": (C.c_int, [C.c_void_p, _dp]),
    "mpm_bin_positions_noalloc": (C.c_int, [C.POINTER(GridDesc), C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p,
                                            C.c_void_p]),
    "mpm_bandwidth_probe": (C.c_int, [C.c_int64, C.c_int32, _dp]),
    "mpm_stream_mix_probe": (C.c_int, [C.c_int64, C.c_int32, _dp]),
    "mpm_set_shifting": (C.c_int, [C.c_void_p, C.c_int32, _ip]),
    "mpm_download_nodal_volume": (C.c_int, [C.c_void_p, _dp]),
    "mpm_download_affine": (C.c_int, [C.c_void_p, _dp, _dp]),
    "mpm_upload_affine": (C.c_int, [C.c_void_p, _dp, _dp]),
    "mpm_upload_accel": (C.c_int, [C.c_void_p, _dp]),
    "mpm_download_accel": (C.c_int, [C.c_void_p, _dp]),
    "mpm_implicit_p2g": (C.c_int, [C.c_void_p, C.c_double]),
    "mpm_implicit_g2p": (C.c_int, [C.c_void_p, _dp, _dp, C.c_double]),
    "mpm_set_fbar": (C.c_int, [C.c_void_p, C.c_int32]),
    "mpm_download_fbar": (C.c_int, [C.c_void_p, _dp, _dp]),
    "mpm_upload_fbar": (C.c_int, [C.c_void_p, _dp, _dp]),
    # include/mpm/mpm_surrogate.h
    "mpm_mlp_forward": (C.c_int, [_fp, _fp, _fp, C.c_int64, _fp]),
    "mpm_mlp_gradient": (C.c_int, [_fp, _fp, _fp, _fp, C.c_int64, _fp, _fp]),
    "mpm_mlp_train": (C.c_int, [_fp, _fp, _fp, _fp, C.c_int64, C.POINTER(TrainParams), _fp]),
    "mpm_surrogate_collect": (C.c_int, [C.c_void_p, C.c_int32, C.c_int64, C.c_double]),
    "mpm_surrogate_fit": (C.c_int, [C.c_void_p, C.POINTER(TrainParams), _fp]),
    "mpm_surrogate_enable": (C.c_int, [C.c_void_p, C.c_int32, C.c_double]),
    "mpm_surrogate_stats_get": (C.c_int, [C.c_void_p, C.POINTER(SurrogateStats)]),
    "mpm_surrogate_reset_counters": (C.c_int, [C.c_void_p]),
    "mpm_surrogate_get_params": (C.c_int, [C.c_void_p, _fp, _fp]),
    "mpm_surrogate_set_params": (C.c_int, [C.c_void_p, _fp, _fp]),
    "mpm_surrogate_download_samples": (C.c_int, [C.c_void_p, _fp, _fp, C.c_int64, C.POINTER(C.c_int64)]),
    # include/mpm/mpm_ensemble.h
    "mpm_ensemble_create": (C.c_int, [C.POINTER(C.c_void_p), C.c_int32, C.POINTER(C.c_int32), C.POINTER(C.c_void_p)]),
    "mpm_ensemble_destroy": (C.c_int, [C.c_void_p]),
    "mpm_ensemble_step": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, _dp]),
    "mpm_ensemble_capture": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32]),
    "mpm_ensemble_launch": (C.c_int, [C.c_void_p, C.c_int32, _dp]),
    "mpm_ensemble_autotune": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.POINTER(TuneResult), C.c_int32,
                                        C.POINTER(C.c_int32)]),
    # include/mpm/mpm_case.h
    "mpm_case_read": (C.c_int, [C.c_char_p, C.c_char_p, C.c_char_p, C.POINTER(C.c_void_p)]),
    "mpm_case_describe": (C.c_int, [C.c_void_p, C.POINTER(CaseDesc)]),
    "mpm_case_materials": (C.c_int, [C.c_void_p, C.POINTER(MaterialDesc), C.c_int32]),
    "mpm_case_particles": (C.c_int, [C.c_void_p, C.POINTER(ParticlesView)]),
    "mpm_case_bc": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(_ip), C.POINTER(_dp), C.POINTER(C.c_int32)]),
    "mpm_case_write": (C.c_int, [C.c_void_p, C.c_char_p, C.c_char_p]),
    "mpm_case_engine": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.POINTER(C.c_void_p)]),
    "mpm_case_free": (None, [C.c_void_p]),
    "mpm_restart_save": (C.c_int, [C.c_char_p, C.POINTER(ParticlesView), C.c_int32, _dp, _dp, _ip, _ip, C.c_int32]),
    "mpm_restart_load": (C.c_int, [C.c_char_p, C.c_int32, C.POINTER(C.c_void_p)]),
    "mpm_restart_view": (C.c_int, [C.c_void_p, C.POINTER(ParticlesView), C.POINTER(_dp), C.POINTER(_dp),
                                   C.POINTER(_ip), C.POINTER(_ip)]),
    "mpm_restart_free": (None, [C.c_void_p]),
    "mpm_restart_write": (C.c_int, [C.c_void_p, C.c_char_p, C.c_int32]),
    "mpm_restart_read": (C.c_int, [C.c_void_p, C.c_char_p]),
    "mpm_get_engine_info": (C.c_int, [C.c_void_p, C.POINTER(EngineInfo)]),
}

MLP_NPARAMS = 6288

_lib = None


class MPMError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"mpm error {code}: {msg}")
        self.code = code


def load(path=None):
    """Loads the engine library (no fallback)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise FileNotFoundError(f"{p} is not built; run hpc-mpm-code_amd/build.py")
    lib = C.CDLL(p)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib


def check(rc):
    if rc != MPM_OK:
        raise MPMError(rc, load().mpm_last_error().decode())
    return rc


def _ptr(a, ct=C.c_double):
    if a is None:
        return None
    return a.ctypes.data_as(C.POINTER(ct))


def grid_desc(xmin, xmax, nelem, idimc=(1, 1, 1)):
    g = GridDesc()
    for i in range(3):
        g.xmin[i] = xmin[i]
        g.xmax[i] = xmax[i]
        g.nelem[i] = int(nelem[i])
        g.idimc[i] = int(idimc[i])
    return g


class Particles:
    """Host particle arrays in the reference AoS FP64 layout."""

    FIELDS = ("coord", "vel", "mass", "vol", "vol0", "stress", "def_grad", "det_def_grad", "trac_force", "matid",
              "eps_p")

    def __init__(self, n, coord=None, vel=None, mass=None, vol0=None, vol=None, stress=None, def_grad=None,
                 det_def_grad=None, trac_force=None, matid=None, eps_p=None):
        self.n = int(n)
        z = lambda *s: np.zeros(s, dtype=np.float64)
        self.coord = np.ascontiguousarray(coord, dtype=np.float64) if coord is not None else z(n, 3)
        self.vel = np.ascontiguousarray(vel, dtype=np.float64) if vel is not None else z(n, 3)
        self.mass = np.ascontiguousarray(mass, dtype=np.float64) if mass is not None else z(n)
        self.vol0 = np.ascontiguousarray(vol0, dtype=np.float64) if vol0 is not None else z(n)
        self.vol = np.ascontiguousarray(vol, dtype=np.float64) if vol is not None else self.vol0.copy()
        self.stress = np.ascontiguousarray(stress, dtype=np.float64) if stress is not None else z(n, 6)
        if def_grad is None:
            def_grad = np.tile(np.eye(3), (n, 1, 1))
        self.def_grad = np.ascontiguousarray(def_grad, dtype=np.float64).reshape(n, 3, 3)
        self.det_def_grad = (np.ascontiguousarray(det_def_grad, dtype=np.float64) if det_def_grad is not None
                             else np.ones(n))
        self.trac_force = (np.ascontiguousarray(trac_force, dtype=np.float64) if trac_force is not None else None)
        self.matid = np.ascontiguousarray(matid, dtype=np.int32) if matid is not None else None
        self.eps_p = np.ascontiguousarray(eps_p, dtype=np.float64) if eps_p is not None else z(n)

    def copy(self):
        p = Particles(self.n)
        for f in self.FIELDS:
            v = getattr(self, f)
            setattr(p, f, None if v is None else v.copy())
        return p

    def view(self):
        v = ParticlesView()
        v.n = self.n
        v.coord = _ptr(self.coord)
        v.vel = _ptr(self.vel)
        v.mass = _ptr(self.mass)
        v.vol = _ptr(self.vol)
        v.vol0 = _ptr(self.vol0)
        v.stress = _ptr(self.stress)
        v.def_grad = _ptr(self.def_grad)
        v.det_def_grad = _ptr(self.det_def_grad)
        v.trac_force = _ptr(self.trac_force)
        v.matid = _ptr(self.matid, C.c_int32)
        v.eps_p = _ptr(self.eps_p)
        return v


class Nodes:
    def __init__(self, nodec):
        self.nodec = int(nodec)
        self.nmass = np.zeros(nodec)
        self.nvof = np.zeros(nodec)
        self.nmome = np.zeros(3 * nodec)
        self.nforce = np.zeros(3 * nodec)
        self.nvel = np.zeros(3 * nodec)
        self.naccel = np.zeros(3 * nodec)

    def view(self):
        v = NodesView()
        v.nodec = self.nodec
        for f in ("nmass", "nvof", "nmome", "nforce", "nvel", "naccel"):
            setattr(v, f, _ptr(getattr(self, f)))
        return v


class Engine:
    """Thin owner of an mpm_engine handle."""

    def __init__(self, scene, precision=PREC_FP64_EXACT, relayout_every=0):
        lib = load()
        self.lib = lib
        self.scene = scene
        idimc = tuple(getattr(scene, "idimc", (1, 1, 1)))
        g = grid_desc(scene.xmin, scene.xmax, scene.nelem, idimc)
        self.nodec = int(np.prod(np.asarray(scene.nelem) + np.asarray(idimc)))
        self.nelem = int(np.prod(scene.nelem))
        mats = (MaterialDesc * len(scene.materials))()
        for i, (model, props) in enumerate(scene.materials):
            mats[i].model = model
            mats[i].nprop = len(props)
            for k, v in enumerate(props):
                mats[i].prop[k] = v
        p = StepParams()
        p.dt = scene.dt
        p.mtol = scene.mtol
        for i in range(3):
            p.bb[i] = scene.bb[i]
        p.scheme = scene.scheme
        p.nlstep = scene.nlstep
        p.precision = precision
        p.relayout_every = relayout_every
        h = C.c_void_p()
        check(lib.mpm_engine_create(C.byref(g), mats, len(scene.materials), C.byref(p), C.byref(h)))
        self.h = h
        for comp, (nbc, fbc) in scene.bcs.items():
            nbc = np.ascontiguousarray(nbc, dtype=np.int32)
            fbc = np.ascontiguousarray(fbc, dtype=np.float64)
            check(lib.mpm_set_bc(self.h, comp, _ptr(nbc, C.c_int32), _ptr(fbc), len(nbc)))
        if getattr(scene, "shifting", False):
            self.set_shifting(True, scene.npxye)
        if getattr(scene, "fbar", False):
            self.set_fbar(True)
        if getattr(scene, "mat_dt", None) is not None:
            self.set_material_dt(scene.mat_dt)

    def set_shifting(self, on, npxye=(2, 2, 2)):
        npx = np.ascontiguousarray(npxye, dtype=np.int32)
        check(self.lib.mpm_set_shifting(self.h, int(on), _ptr(npx, C.c_int32)))

    def affine(self, n, dinv=False):
        """TPIC velocity gradient / APIC B matrix [n][3][3] (and APIC inverse D)."""
        cm = np.zeros((n, 3, 3))
        di = np.zeros((n, 3, 3)) if dinv else None
        check(self.lib.mpm_download_affine(self.h, _ptr(cm), _ptr(di)))
        return (cm, di) if dinv else cm

    def set_affine(self, cmat, dinv=None):
        cm = np.ascontiguousarray(cmat, np.float64)
        di = None if dinv is None else np.ascontiguousarray(dinv, np.float64)
        check(self.lib.mpm_upload_affine(self.h, _ptr(cm), _ptr(di)))

    def accel(self, n):
        """Particle accelerations of the implicit transfers, [n][3]."""
        a = np.zeros((n, 3))
        check(self.lib.mpm_download_accel(self.h, _ptr(a)))
        return a

    def set_accel(self, accel):
        a = np.ascontiguousarray(accel, np.float64)
        check(self.lib.mpm_upload_accel(self.h, _ptr(a)))

    def implicit_p2g(self, facl=1.0):
        check(self.lib.mpm_implicit_p2g(self.h, facl))

    def implicit_g2p(self, ndispl, nb_para, gamma_nb):
        nd = np.ascontiguousarray(ndispl, np.float64)
        nb = np.ascontiguousarray(nb_para, np.float64)
        assert nd.size == 3 * self.nodec and nb.size == 6
        check(self.lib.mpm_implicit_g2p(self.h, _ptr(nd), _ptr(nb), gamma_nb))

    def info(self):
        i = EngineInfo()
        check(self.lib.mpm_get_engine_info(self.h, C.byref(i)))
        return {k: getattr(i, k) for k, _ in EngineInfo._fields_}

    def restart_write(self, path, digits=0):
        check(self.lib.mpm_restart_write(self.h, _path(path), digits))

    def restart_read(self, path):
        check(self.lib.mpm_restart_read(self.h, _path(path)))

    def set_fbar(self, on=True):
        check(self.lib.mpm_set_fbar(self.h, int(on)))

    def fbar_state(self, n):
        """(Fbar [n][3][3], Jbar [n]) of the F-bar projection."""
        fb = np.zeros((n, 3, 3))
        jb = np.zeros(n)
        check(self.lib.mpm_download_fbar(self.h, _ptr(fb), _ptr(jb)))
        return fb, jb

    def set_fbar_state(self, fbar, jbar):
        check(self.lib.mpm_upload_fbar(self.h, _ptr(np.ascontiguousarray(fbar, np.float64)),
                                       _ptr(np.ascontiguousarray(jbar, np.float64))))

    def nodal_volume(self):
        out = np.zeros(self.nodec)
        check(self.lib.mpm_download_nodal_volume(self.h, _ptr(out)))
        return out

    def close(self):
        if getattr(self, "h", None):
            self.lib.mpm_engine_destroy(self.h)
            self.h = None

    __del__ = close

    def upload(self, parts):
        check(self.lib.mpm_upload_particles(self.h, C.byref(parts.view())))

    def download(self, n):
        out = Particles(n)
        check(self.lib.mpm_download_particles(self.h, C.byref(out.view())))
        return out

    def nodes(self):
        nd = Nodes(self.nodec)
        check(self.lib.mpm_download_nodes(self.h, C.byref(nd.view())))
        return nd

    def bin(self):
        check(self.lib.mpm_bin(self.h))

    def p2g(self, facl=1.0):
        check(self.lib.mpm_p2g(self.h, facl))

    def solve_grid(self):
        check(self.lib.mpm_solve_grid(self.h))

    def g2p(self, facl=1.0):
        check(self.lib.mpm_g2p(self.h, facl))

    def step(self, istep0, nsteps):
        return self.lib.mpm_step(self.h, istep0, nsteps)

    def bins(self, n):
        idepf = np.zeros(self.nelem, np.int32)
        idp2p = np.zeros(n, np.int32)
        numep = np.zeros(self.nelem, np.int32)
        cell = np.zeros(n, np.int32)
        check(self.lib.mpm_download_bins(self.h, _ptr(idepf, C.c_int32), _ptr(idp2p, C.c_int32),
                                         _ptr(numep, C.c_int32), _ptr(cell, C.c_int32)))
        return idepf, idp2p, numep, cell

    def set_timing(self, on=True):
        check(self.lib.mpm_set_timing(self.h, 1 if on else 0))

    def set_material_dt(self, dt):
        """Per-material time steps (None restores the common dt)."""
        d = None if dt is None else np.ascontiguousarray(dt, np.float64)
        check(self.lib.mpm_set_material_dt(self.h, _ptr(d)))

    def set_option(self, key, value):
        check(self.lib.mpm_set_option(self.h, key, value))

    def kernel_times(self):
        kt = KernelTimes()
        check(self.lib.mpm_get_kernel_times(self.h, C.byref(kt)))
        return {f: getattr(kt, f) for f, _ in KernelTimes._fields_}


    # ---- J2 constitutive surrogate
    def surrogate_collect(self, on=True, capacity=1 << 20, rate=1.0):
        check(self.lib.mpm_surrogate_collect(self.h, int(on), int(capacity), float(rate)))

    def surrogate_fit(self, epochs=20, batch=8192, lr=2e-3, lr_final=5e-5, seed=1):
        tp = TrainParams(epochs, batch, lr, lr_final, seed, 0)
        loss = C.c_float(0)
        check(self.lib.mpm_surrogate_fit(self.h, C.byref(tp), C.byref(loss)))
        return loss.value

    def surrogate_enable(self, on=True, tol=1e-4):
        check(self.lib.mpm_surrogate_enable(self.h, int(on), float(tol)))

    def surrogate_stats(self):
        st = SurrogateStats()
        check(self.lib.mpm_surrogate_stats_get(self.h, C.byref(st)))
        return {k: getattr(st, k) for k, _ in SurrogateStats._fields_}

    def surrogate_reset_counters(self):
        check(self.lib.mpm_surrogate_reset_counters(self.h))

    def surrogate_params(self):
        p = np.zeros(MLP_NPARAMS, np.float32)
        sc = np.zeros(64, np.float32)
        check(self.lib.mpm_surrogate_get_params(self.h, _ptr(p, C.c_float), _ptr(sc, C.c_float)))
        return p, sc

    def surrogate_set_params(self, params, scaling):
        p = np.ascontiguousarray(params, np.float32)
        sc = np.ascontiguousarray(scaling, np.float32)
        check(self.lib.mpm_surrogate_set_params(self.h, _ptr(p, C.c_float), _ptr(sc, C.c_float)))

    def surrogate_samples(self, max_rows=1 << 20):
        xin = np.zeros((max_rows, 16), np.float32)
        tgt = np.zeros((max_rows, 16), np.float32)
        n = C.c_int64(0)
        check(self.lib.mpm_surrogate_download_samples(self.h, _ptr(xin, C.c_float), _ptr(tgt, C.c_float), max_rows,
                                                      C.byref(n)))
        return xin[:n.value], tgt[:n.value]


class Ensemble:
    """Independent scenes advanced together (include/mpm/mpm_ensemble.h). The
    engines stay owned by this object and are destroyed with it."""

    def __init__(self, scene_list, bucket=None, precision=PREC_FP32, relayout_every=16):
        self.lib = load()
        self.engines = []
        for sc in scene_list:
            e = Engine(sc, precision=precision, relayout_every=relayout_every if precision == PREC_FP32 else 0)
            e.upload(sc.parts)
            if precision == PREC_FP32:
                e.set_option(OPT_FULL_NODE_OUTPUT, 0)
            self.engines.append(e)
        arr = (C.c_void_p * len(self.engines))(*[e.h.value for e in self.engines])
        b = None if bucket is None else (C.c_int32 * len(bucket))(*bucket)
        h = C.c_void_p()
        check(self.lib.mpm_ensemble_create(arr, len(self.engines), b, C.byref(h)))
        self.h = h

    def step(self, istep0, nsteps):
        ms = C.c_double(0)
        check(self.lib.mpm_ensemble_step(self.h, istep0, nsteps, C.byref(ms)))
        return ms.value

    def capture(self, istep0, block_steps):
        check(self.lib.mpm_ensemble_capture(self.h, istep0, block_steps))

    def launch(self, nlaunch):
        ms = C.c_double(0)
        check(self.lib.mpm_ensemble_launch(self.h, nlaunch, C.byref(ms)))
        return ms.value

    def autotune(self, istep0, trial_steps=5):
        res = (TuneResult * 16)()
        nres = C.c_int32(0)
        check(self.lib.mpm_ensemble_autotune(self.h, istep0, trial_steps, res, 16, C.byref(nres)))
        return [{k: getattr(res[i], k) for k, _ in TuneResult._fields_} for i in range(nres.value)]

    def close(self):
        if getattr(self, "h", None):
            self.lib.mpm_ensemble_destroy(self.h)
            self.h = None
        for e in self.engines:
            e.close()
        self.engines = []

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _path(p):
    return None if p is None else os.fsencode(p)


class Case:
    """A reference solid case read from the generator's files
    (include/mpm/mpm_case.h): input.txt parameter block, griddata.txt,
    spdata.txt. Any of the three may be None."""

    def __init__(self, input_txt=None, griddata_txt=None, spdata_txt=None):
        self.lib = load()
        h = C.c_void_p()
        check(self.lib.mpm_case_read(_path(input_txt), _path(griddata_txt), _path(spdata_txt), C.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.lib.mpm_case_free(self.h)
            self.h = None

    __del__ = close

    def desc(self):
        d = CaseDesc()
        check(self.lib.mpm_case_describe(self.h, C.byref(d)))
        return d

    def materials(self):
        d = self.desc()
        m = (MaterialDesc * max(d.nmat, 1))()
        check(self.lib.mpm_case_materials(self.h, m, d.nmat))
        return [(m[i].model, [m[i].prop[k] for k in range(m[i].nprop)]) for i in range(d.nmat)]

    def particles(self):
        v = ParticlesView()
        check(self.lib.mpm_case_particles(self.h, C.byref(v)))
        n = int(v.n)
        arr = lambda ptr, shape: np.ctypeslib.as_array(ptr, shape=shape).copy() if n else np.zeros(shape)
        matid = np.ctypeslib.as_array(v.matid, shape=(n,)).copy() if n else np.zeros(0, np.int32)
        return Particles(n, coord=arr(v.coord, (n, 3)), mass=arr(v.mass, (n,)), vol0=arr(v.vol0, (n,)),
                         matid=matid)

    def bcs(self):
        out = {}
        for comp in range(3):
            nbc, fbc, ibc = _ip(), _dp(), C.c_int32()
            check(self.lib.mpm_case_bc(self.h, comp, C.byref(nbc), C.byref(fbc), C.byref(ibc)))
            k = ibc.value
            out[comp] = ((np.ctypeslib.as_array(nbc, shape=(k,)).copy(), np.ctypeslib.as_array(fbc, shape=(k,)).copy())
                         if k else (np.zeros(0, np.int32), np.zeros(0)))
        return out

    def write(self, griddata_txt=None, spdata_txt=None):
        check(self.lib.mpm_case_write(self.h, _path(griddata_txt), _path(spdata_txt)))

    def engine(self, precision=PREC_FP64_EXACT, relayout_every=0, shifting=False):
        """Engine built by mpm_case_engine (grid, materials, BCs, points
        uploaded; reference particle shifting on when asked)."""
        h = C.c_void_p()
        check(self.lib.mpm_case_engine(self.h, precision, relayout_every, int(shifting), C.byref(h)))
        d = self.desc()
        e = Engine.__new__(Engine)
        e.lib, e.h, e.scene = self.lib, h, None
        e.nodec = int(np.prod([d.grid.nelem[a] + d.grid.idimc[a] for a in range(3)]))
        e.nelem = int(np.prod([d.grid.nelem[a] for a in range(3)]))
        return e

    def scene(self, name="case"):
        """The case as a scene description (the form the oracle and the
        Python-side engine constructor take)."""
        from . import scenes
        d = self.desc()
        g = d.grid
        s = scenes.Scene(name, tuple(g.xmin), tuple(g.xmax), tuple(g.nelem), self.materials(), d.rho, d.step.dt,
                         max(d.iend - d.ista + 1, 0), tuple(d.step.bb), mtol=d.step.mtol, scheme=d.step.scheme,
                         nlstep=d.step.nlstep, bcs=self.bcs(), parts=self.particles(), idimc=tuple(g.idimc),
                         npxye=tuple(d.npxye))
        return s


def restart_save(path, parts, scheme=SCHEME_PIC, cmat=None, dinv=None, ids=None, surf=None, digits=0):
    """Writes a reference restart file (include/mpm/mpm_case.h) from host arrays."""
    lib = load()
    ids = None if ids is None else np.ascontiguousarray(ids, np.int32)
    surf = None if surf is None else np.ascontiguousarray(surf, np.int32)
    cm = None if cmat is None else np.ascontiguousarray(cmat, np.float64)
    di = None if dinv is None else np.ascontiguousarray(dinv, np.float64)
    check(lib.mpm_restart_save(_path(path), C.byref(parts.view()), scheme, _ptr(cm), _ptr(di),
                               _ptr(ids, C.c_int32), _ptr(surf, C.c_int32), digits))


def restart_load(path, scheme=SCHEME_PIC):
    """Reads a reference restart file: (Particles, cmat or None, dinv or None, ids, surf)."""
    lib = load()
    h = C.c_void_p()
    check(lib.mpm_restart_load(_path(path), scheme, C.byref(h)))
    try:
        v = ParticlesView()
        cm, di, ids, surf = _dp(), _dp(), _ip(), _ip()
        check(lib.mpm_restart_view(h, C.byref(v), C.byref(cm), C.byref(di), C.byref(ids), C.byref(surf)))
        n = int(v.n)
        a = lambda ptr, shape: np.ctypeslib.as_array(ptr, shape=shape).copy() if n and ptr else None
        parts = Particles(n, coord=a(v.coord, (n, 3)), vel=a(v.vel, (n, 3)), mass=a(v.mass, (n,)),
                          vol0=a(v.vol0, (n,)), vol=a(v.vol, (n,)), stress=a(v.stress, (n, 6)),
                          def_grad=a(v.def_grad, (n, 3, 3)), det_def_grad=a(v.det_def_grad, (n,)),
                          matid=a(v.matid, (n,)))
        return parts, a(cm, (n, 3, 3)), a(di, (n, 3, 3)), a(ids, (n,)), a(surf, (n,))
    finally:
        lib.mpm_restart_free(h)


def mlp_forward(params, scaling, inp):
    lib = load()
    p = np.ascontiguousarray(params, np.float32)
    sc = np.ascontiguousarray(scaling, np.float32)
    x = np.ascontiguousarray(inp, np.float32)
    out = np.zeros((x.shape[0], 16), np.float32)
    check(lib.mpm_mlp_forward(_ptr(p, C.c_float), _ptr(sc, C.c_float), _ptr(x, C.c_float), x.shape[0],
                              _ptr(out, C.c_float)))
    return out


def mlp_gradient(params, scaling, inp, tgt):
    lib = load()
    p = np.ascontiguousarray(params, np.float32)
    sc = np.ascontiguousarray(scaling, np.float32)
    x = np.ascontiguousarray(inp, np.float32)
    t = np.ascontiguousarray(tgt, np.float32)
    g = np.zeros(MLP_NPARAMS, np.float32)
    loss = C.c_float(0)
    check(lib.mpm_mlp_gradient(_ptr(p, C.c_float), _ptr(sc, C.c_float), _ptr(x, C.c_float), _ptr(t, C.c_float),
                               x.shape[0], _ptr(g, C.c_float), C.byref(loss)))
    return loss.value, g


def mlp_train(params, scaling, inp, tgt, epochs=10, batch=4096, lr=2e-3, lr_final=5e-5, seed=1):
    lib = load()
    p = np.array(params, np.float32)
    sc = np.ascontiguousarray(scaling, np.float32)
    x = np.ascontiguousarray(inp, np.float32)
    t = np.ascontiguousarray(tgt, np.float32)
    hist = np.zeros(epochs, np.float32)
    tp = TrainParams(epochs, batch, lr, lr_final, seed, 0)
    check(lib.mpm_mlp_train(_ptr(p, C.c_float), _ptr(sc, C.c_float), _ptr(x, C.c_float), _ptr(t, C.c_float),
                            x.shape[0], C.byref(tp), _ptr(hist, C.c_float)))
    return p, hist


def maksf_batch(grid, cells, xyg):
    lib = load()
    cells = np.ascontiguousarray(cells, dtype=np.int32)
    xyg = np.ascontiguousarray(xyg, dtype=np.float64)
    n = len(cells)
    nen = int(np.prod(np.asarray(list(grid.idimc)) + 1))
    nc = np.zeros((n, nen), np.int32)
    sf = np.zeros((n, nen))
    dsf = np.zeros((n, nen, 3))
    check(lib.mpm_maksf_batch(C.byref(grid), n, _ptr(cells, C.c_int32), _ptr(xyg), _ptr(nc, C.c_int32), _ptr(sf),
                              _ptr(dsf)))
    return nc, sf, dsf


def bin_positions(grid, xyg):
    lib = load()
    xyg = np.ascontiguousarray(xyg, dtype=np.float64)
    n = len(xyg)
    cell = np.zeros(n, np.int32)
    order = np.zeros(n, np.int32)
    check(lib.mpm_bin_positions(C.byref(grid), n, _ptr(xyg), _ptr(cell, C.c_int32), _ptr(order, C.c_int32)))
    return cell, order


def bandwidth_probe(nbytes=1 << 30, iters=20):
    """Achieved copy bandwidth in GB/s (read + write) of a plain copy kernel."""
    out = C.c_double()
    check(load().mpm_bandwidth_probe(int(nbytes), int(iters), C.byref(out)))
    return out.value


def stream_mix_probe(n=1 << 22, iters=20):
    """Achieved GB/s (read + write) of the particle update's stream mix over
    n particles (15 fields read, 19 written, structure of arrays)."""
    out = C.c_double()
    check(load().mpm_stream_mix_probe(int(n), int(iters), C.byref(out)))
    return out.value
