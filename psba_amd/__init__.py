"""MI355X bundle adjuster: Python view of the C-ABI in include/psba_capi.h.

The compute path is libpsba.so (hand-written HIP for gfx950, built by
``python -m psba_amd.build``).  This module only marshals numpy arrays through
ctypes; it has no compute fallback: if the library is missing, importing the
solver classes raises.

Mirrors the reference's entry points (PSBA/levmar.h, PSBA/trust_region.h,
PSBA/cl_psba.h): ``Context.levmar`` / ``Context.trust_region`` /
``Context.hybrid`` with the shared iteration counter ``itno`` and the ITER_*
return codes of PSBA/psba.h.
"""
import ctypes as ct
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpsba.so")

# PSBA/psba.h:12-18
ITER_TURN_TO_LM = 1
ITER_TURN_TO_TR = 2
ITER_CONTINUE = 3
ITER_ERR = 4
ITER_DP_NO_CHANGE = 5
ITER_ERR_SMALL_ENOUGH = 6
ITER_PASS = 7

PREC_FP64 = 0
PREC_MIXED = 1

LOSS_NONE, LOSS_HUBER, LOSS_CAUCHY = 0, 1, 2

BUF = dict(EX=0, JA=1, JB=2, U=3, V=4, W=5, G=6, VINV=7, S=9, EA=10, DP=11, EB=12, L=13)
INDEX = dict(OBS_PT=0, OBS_CAM=1, PT_PTR=2, CAM_PTR=3, CAM_OBS=4, CCHUNK_BEGIN=5, CCHUNK_END=6, CAM_CCHUNK_PTR=7,
             PAIR_ITEMS=8, PAIR_K=9, PAIR_L=10, SCHUNK_BEGIN=11, SCHUNK_END=12, PAIR_SCHUNK_PTR=13, LBLOCK_PAIR=14,
             SCHUNK_ORDER=15, PT_ORDER=16)


class PsbaError(RuntimeError):
    pass


class Desc(ct.Structure):
    _fields_ = [("cnp", ct.c_int), ("pnp", ct.c_int), ("mnp", ct.c_int), ("n_cams", ct.c_int),
                ("n_pts", ct.c_int), ("n_obs", ct.c_int), ("precision", ct.c_int), ("device", ct.c_int)]


FLAG_NO_HANDOFF = 1
FLAG_HOST_CONTROL = 2  # psba_levmar decides each trial on the host


class Opts(ct.Structure):
    _fields_ = [("max_iter", ct.c_int), ("max_pb_tries", ct.c_int), ("refine_steps", ct.c_int),
                ("flags", ct.c_int), ("rel_tol", ct.c_double)]


class TraceRec(ct.Structure):
    _fields_ = [("itno", ct.c_int), ("kind", ct.c_int), ("factor_ok", ct.c_int), ("accepted", ct.c_int),
                ("cost", ct.c_double), ("damp", ct.c_double), ("rho", ct.c_double), ("lambda_", ct.c_double),
                ("dp_sq", ct.c_double)]


class Timing(ct.Structure):
    _fields_ = [("jacobian_normal", ct.c_double), ("schur", ct.c_double), ("factor", ct.c_double),
                ("solve", ct.c_double), ("backsub", ct.c_double), ("residual", ct.c_double),
                ("other", ct.c_double), ("trials", ct.c_int), ("jacobians", ct.c_int)]


class Info(ct.Structure):
    _fields_ = [("cnp", ct.c_int), ("n_cams", ct.c_int), ("n_pts", ct.c_int), ("n_obs", ct.c_int), ("n", ct.c_int),
                ("n_pad", ct.c_int), ("n_pairs", ct.c_int), ("n_pair_items", ct.c_long), ("sum_t2", ct.c_double)]


PROBE_OFF, PROBE_SCHUR, PROBE_FACTOR, PROBE_JACOBIAN, PROBE_BSOLVE = 0, 1, 2, 3, 4


class _SceneC(ct.Structure):
    _fields_ = [("cnp", ct.c_int), ("n_cams", ct.c_int), ("n_pts", ct.c_int), ("n_obs", ct.c_int),
                ("K", ct.POINTER(ct.c_double)), ("q0", ct.POINTER(ct.c_double)),
                ("cams", ct.POINTER(ct.c_double)), ("pts", ct.POINTER(ct.c_double)),
                ("meas", ct.POINTER(ct.c_double)), ("iidx", ct.POINTER(ct.c_int)),
                ("jidx", ct.POINTER(ct.c_int)), ("fixmask", ct.POINTER(ct.c_ubyte)),
                ("gt_q0", ct.POINTER(ct.c_double)), ("gt_cams", ct.POINTER(ct.c_double)),
                ("gt_pts", ct.POINTER(ct.c_double)), ("sum_t2", ct.c_double), ("n_outliers", ct.c_int),
                ("zipf_s", ct.c_double)]


class SceneOpts(ct.Structure):
    _fields_ = [("cnp", ct.c_int), ("n_cams", ct.c_int), ("n_pts", ct.c_int), ("zipf_s", ct.c_double),
                ("noise", ct.c_double), ("outlier_frac", ct.c_double), ("perturb", ct.c_double),
                ("size_class", ct.c_int), ("problem", ct.c_ulonglong)]


class SbaData(ct.Structure):
    _fields_ = [("n_cams", ct.c_int), ("n_pts", ct.c_int), ("n_obs", ct.c_int), ("cnp", ct.c_int),
                ("pnp", ct.c_int), ("mnp", ct.c_int), ("motstruct", ct.POINTER(ct.c_double)),
                ("initrot", ct.POINTER(ct.c_double)), ("imgpts", ct.POINTER(ct.c_double)),
                ("covimgpts", ct.POINTER(ct.c_double)), ("vmask", ct.POINTER(ct.c_char))]


class ReportStages(ct.Structure):
    _fields_ = [("jacobian_normal", ct.c_double), ("schur", ct.c_double), ("factor", ct.c_double),
                ("solve", ct.c_double), ("backsub", ct.c_double), ("residual", ct.c_double),
                ("other", ct.c_double), ("trials", ct.c_int), ("jacobians", ct.c_int)]


class Report(ct.Structure):
    _fields_ = [("code", ct.c_int), ("itno", ct.c_int), ("lm_iterations", ct.c_int), ("tr_iterations", ct.c_int),
                ("handoffs", ct.c_int), ("lm_trials", ct.c_int), ("tr_steps", ct.c_int),
                ("init_sse", ct.c_double), ("final_sse", ct.c_double), ("init_error_ref", ct.c_double),
                ("final_error_ref", ct.c_double), ("init_rms_px", ct.c_double), ("final_rms_px", ct.c_double),
                ("wall_ms", ct.c_double), ("stages", ReportStages)]


class BatchDesc(ct.Structure):
    _fields_ = [("cnp", ct.c_int), ("precision", ct.c_int), ("device", ct.c_int), ("max_problems", ct.c_int)]


class ServerDesc(ct.Structure):
    _fields_ = [("cnp", ct.c_int), ("precision", ct.c_int), ("device", ct.c_int), ("lanes", ct.c_int),
                ("max_batch", ct.c_int), ("hbm_budget", ct.c_longlong), ("max_iter", ct.c_int),
                ("rel_tol", ct.c_double)]


class ServerStats(ct.Structure):
    _fields_ = [("submitted", ct.c_longlong), ("completed", ct.c_longlong), ("rejected", ct.c_longlong),
                ("batches", ct.c_longlong), ("bytes_held", ct.c_longlong), ("bytes_peak", ct.c_longlong),
                ("queued", ct.c_int), ("running", ct.c_int)]


_lib = None


def lib():
    """Load libpsba.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise PsbaError("libpsba.so not built; run `python -m psba_amd.build`")
    L = ct.CDLL(LIB_PATH)
    vp, dp, ip = ct.c_void_p, ct.POINTER(ct.c_double), ct.POINTER(ct.c_int)
    sig = {
        "psba_create": (ct.c_int, [ct.POINTER(Desc), ct.POINTER(vp)]),
        "psba_destroy": (ct.c_int, [vp]),
        "psba_upload": (ct.c_int, [vp, dp, dp, dp, dp, dp]),
        "psba_set_index": (ct.c_int, [vp, ip, ip]),
        "psba_set_fixed": (ct.c_int, [vp, ct.POINTER(ct.c_ubyte)]),
        "psba_read_params": (ct.c_int, [vp, dp, dp]),
        "psba_write_params": (ct.c_int, [vp, dp, dp]),
        "psba_snapshot_params": (ct.c_int, [vp]),
        "psba_restore_params": (ct.c_int, [vp]),
        "psba_strerror": (ct.c_char_p, [ct.c_int]),
        "psba_last_error": (ct.c_char_p, []),
        "psba_device_count": (ct.c_int, []),
        "psba_device_info": (ct.c_int, [ct.c_int, ct.c_char_p, ct.c_int, ct.POINTER(ct.c_int), ct.POINTER(ct.c_int),
                                        ct.POINTER(ct.c_longlong)]),
        "psba_generate_idxs": (ct.c_int, [ct.c_int, ct.c_int, ct.c_int, ct.c_char_p, ip, ip, ip, ip, ip]),
        "psba_levmar": (ct.c_int, [vp, ct.POINTER(Opts), ip, dp, ct.POINTER(TraceRec), ct.c_int, ip]),
        "psba_trust_region": (ct.c_int, [vp, ct.POINTER(Opts), ip, dp, ct.POINTER(TraceRec), ct.c_int, ip]),
        "psba_hybrid": (ct.c_int, [vp, ct.POINTER(Opts), ip, dp, dp, ct.POINTER(TraceRec), ct.c_int, ip]),
        "psba_stage_residual": (ct.c_int, [vp, ct.c_int, dp]),
        "psba_stage_jacobian_normal": (ct.c_int, [vp, ct.c_double, ct.c_double]),
        "psba_stage_schur": (ct.c_int, [vp, ct.c_double]),
        "psba_stage_factor_solve": (ct.c_int, [vp, ip]),
        "psba_stage_backsub": (ct.c_int, [vp]),
        "psba_stage_modchol": (ct.c_int, [vp, dp]),
        "psba_stage_refine": (ct.c_int, [vp, ct.c_double, ct.c_int]),
        "psba_stage_load_S": (ct.c_int, [vp, dp, dp]),
        "psba_stage_vinv": (ct.c_int, [vp, dp, ct.c_double, dp]),
        "psba_buffer_size": (ct.c_long, [vp, ct.c_int]),
        "psba_read_buffer": (ct.c_int, [vp, ct.c_int, dp, ct.c_long]),
        "psba_get_timing": (ct.c_int, [vp, ct.POINTER(Timing)]),
        "psba_enable_timing": (ct.c_int, [vp, ct.c_int]),
        "psba_get_info": (ct.c_int, [vp, ct.POINTER(Info)]),
        "psba_probe": (ct.c_int, [vp, ct.c_int]),
        "psba_probe_read": (ct.c_int, [vp, dp, ip]),
        "psba_scene_generate": (ct.POINTER(_SceneC), [ct.c_int, ct.c_ulonglong, ct.POINTER(SceneOpts)]),
        "psba_scene_free": (None, [ct.POINTER(_SceneC)]),
        "psba_problem_seed": (ct.c_ulonglong, [ct.c_ulonglong, ct.c_ulonglong]),
        "psba_read_sba": (ct.c_int, [ct.c_char_p, ct.c_char_p, ct.c_int, ct.c_int, ct.c_int, ct.c_void_p, ct.c_int,
                                     ct.POINTER(SbaData)]),
        "psba_sba_free": (None, [ct.POINTER(SbaData)]),
        "psba_quat2vec": (None, [dp, ct.c_int, dp, ct.c_int]),
        "psba_write_sba": (ct.c_int, [ct.c_char_p, ct.c_char_p, ct.c_int, ct.c_int, ct.c_int, dp, dp, dp, dp,
                                      ct.POINTER(SbaData)]),
        "psba_levmar_entry_cost": (ct.c_int, [vp, dp]),
        "psba_problem_bytes": (ct.c_longlong, [ct.c_int, ct.c_int, ct.c_int, ct.c_int, ct.c_int, ct.c_longlong]),
        "psba_batch_create": (ct.c_int, [ct.POINTER(BatchDesc), ct.POINTER(vp)]),
        "psba_batch_destroy": (ct.c_int, [vp]),
        "psba_batch_bytes": (ct.c_longlong, [vp]),
        "psba_batch_reset": (ct.c_int, [vp]),
        "psba_batch_add": (ct.c_int, [vp, ct.c_int, ct.c_int, ct.c_int, dp, dp, dp, dp, dp, ip, ip,
                                      ct.POINTER(ct.c_ubyte), ip]),
        "psba_batch_commit": (ct.c_int, [vp]),
        "psba_batch_launch": (ct.c_int, [vp, ct.POINTER(Opts)]),
        "psba_batch_poll": (ct.c_int, [vp]),
        "psba_batch_solve": (ct.c_int, [vp, ct.POINTER(Opts), ip]),
        "psba_batch_result": (ct.c_int, [vp, ct.c_int, ip, ip, dp, dp, dp, dp]),
        "psba_batch_elapsed": (ct.c_int, [vp, dp, ip]),
        "psba_batch_autotune": (ct.c_int, [vp, ip, ip]),
        "psba_server_create": (ct.c_int, [ct.POINTER(ServerDesc), ct.POINTER(vp)]),
        "psba_server_destroy": (ct.c_int, [vp]),
        "psba_server_clock": (ct.c_double, []),
        "psba_server_submit": (ct.c_int, [vp, ct.c_int, ct.c_int, ct.c_int, dp, dp, dp, dp, dp, ip, ip,
                                          ct.POINTER(ct.c_ubyte), ct.POINTER(ct.c_longlong)]),
        "psba_server_step": (ct.c_int, [vp]),
        "psba_server_drain": (ct.c_int, [vp]),
        "psba_server_result": (ct.c_int, [vp, ct.c_longlong, ip, ip, dp, dp, dp, dp, dp, dp]),
        "psba_server_get_stats": (ct.c_int, [vp, ct.POINTER(ServerStats)]),
        "psba_server_completed": (ct.c_int, [vp, ct.POINTER(ct.c_longlong), ct.c_int]),
        "psba_set_index_device": (ct.c_int, [vp, ip, ip]),
        "psba_set_loss": (ct.c_int, [vp, ct.c_int, ct.c_double]),
        "psba_hybrid_report": (ct.c_int, [vp, ct.POINTER(Opts), ip, ct.POINTER(Report)]),
        "psba_generate_idxs_device": (ct.c_int, [ct.c_int, ct.c_int, ct.c_int, ct.c_int, ct.c_char_p, ip, ip, ip,
                                                 ip]),
        "psba_index_size": (ct.c_long, [vp, ct.c_int]),
        "psba_read_index": (ct.c_int, [vp, ct.c_int, ip, ct.c_long]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def _dptr(a):
    return a.ctypes.data_as(ct.POINTER(ct.c_double))


def _iptr(a):
    return a.ctypes.data_as(ct.POINTER(ct.c_int))


def _check(rc, what):
    if rc < 0:
        L = lib()
        raise PsbaError("%s: %s (%s)" % (what, L.psba_strerror(rc).decode(), L.psba_last_error().decode()))
    return rc


def device_count():
    return lib().psba_device_count()


def device_info(device=0):
    """dict(name, cus, clock_mhz, mem_bytes) of a HIP device."""
    name = ct.create_string_buffer(256)
    cus, clk, mem = ct.c_int(), ct.c_int(), ct.c_longlong()
    _check(lib().psba_device_info(device, name, 256, ct.byref(cus), ct.byref(clk), ct.byref(mem)), "psba_device_info")
    return dict(name=name.value.decode(), cus=cus.value, clock_mhz=clk.value, mem_bytes=mem.value)


def read_sba(cams_path, pts_path, cnp=11, pnp=3, mnp=2, quat_filter=True, filecnp=None):
    """PSBA/readparams.cpp:444-541 through psba_read_sba.  With quat_filter the
    camera lines are passed through quat2vec (as PSBA/main.cpp:102-108 does),
    each holding filecnp = cnp + 1 values.  Returns a dict of numpy copies."""
    L = lib()
    d = SbaData()
    filt = ct.cast(L.psba_quat2vec, ct.c_void_p) if quat_filter else None
    fc = (cnp + 1 if filecnp is None else filecnp) if quat_filter else 0
    _check(L.psba_read_sba(os.fsencode(cams_path), os.fsencode(pts_path), cnp, pnp, mnp, filt, fc, ct.byref(d)),
           "psba_read_sba")
    try:
        nc, npt, no = d.n_cams, d.n_pts, d.n_obs

        def arr(ptr, n):
            return np.ctypeslib.as_array(ptr, shape=(n,)).copy() if n else np.zeros(0)

        ms = arr(d.motstruct, nc * cnp + npt * pnp)
        out = dict(n_cams=nc, n_pts=npt, n_obs=no, cams=ms[:nc * cnp], pts=ms[nc * cnp:],
                   initrot=arr(d.initrot, 4 * nc), imgpts=arr(d.imgpts, no * mnp),
                   cov=arr(d.covimgpts, no * mnp * mnp) if d.covimgpts else None,
                   vmask=np.frombuffer(ct.string_at(d.vmask, npt * nc), np.int8).reshape(npt, nc).copy())
    finally:
        L.psba_sba_free(ct.byref(d))
    return out


def write_sba(cams_path, pts_path, cnp, cams, pts, initrot, imgpts, vmask, K=None, cov=None, mnp=2):
    """Solved parameters in the sba text format through psba_write_sba: each
    camera line holds K (when given), the leading intrinsics, the whole
    rotation (local rotation composed with initrot) as a unit quaternion and
    the translation; point lines repeat the measured projections."""
    L = lib()
    vmask = np.ascontiguousarray(vmask, dtype=np.int8)
    npt, nc = vmask.shape
    imgpts = np.ascontiguousarray(imgpts, dtype=np.float64)
    cams = np.ascontiguousarray(cams, dtype=np.float64)
    pts = np.ascontiguousarray(pts, dtype=np.float64)
    initrot = np.ascontiguousarray(initrot, dtype=np.float64)
    Karr = None if K is None else np.ascontiguousarray(K, dtype=np.float64)
    covarr = None if cov is None else np.ascontiguousarray(cov, dtype=np.float64)
    d = SbaData(n_cams=nc, n_pts=npt, n_obs=imgpts.size // mnp, cnp=cnp, pnp=3, mnp=mnp)
    d.imgpts = _dptr(imgpts)
    d.vmask = vmask.ctypes.data_as(ct.POINTER(ct.c_char))
    if covarr is not None:
        d.covimgpts = _dptr(covarr)
    _check(L.psba_write_sba(os.fsencode(cams_path), os.fsencode(pts_path), nc, npt, cnp,
                            None if Karr is None else _dptr(Karr), _dptr(initrot), _dptr(cams), _dptr(pts),
                            ct.byref(d)), "psba_write_sba")


class Scene:
    """A seeded synthetic scene (include/psba_scene.h); all arrays are numpy copies."""

    def __init__(self, config, seed, **kw):
        o = SceneOpts(noise=-1.0, outlier_frac=-1.0, perturb=-1.0)
        for k, v in kw.items():
            setattr(o, k, v)
        p = lib().psba_scene_generate(int(config), ct.c_ulonglong(seed), ct.byref(o))
        if not p:
            raise PsbaError("scene generation failed")
        s = p.contents
        C, nc, npt, no = s.cnp, s.n_cams, s.n_pts, s.n_obs

        def arr(ptr, n, dt):
            return np.ctypeslib.as_array(ptr, shape=(n,)).astype(dt, copy=True)

        self.config, self.seed = config, seed
        self.cnp, self.n_cams, self.n_pts, self.n_obs = C, nc, npt, no
        self.K = arr(s.K, 5 * nc, np.float64)
        self.q0 = arr(s.q0, 4 * nc, np.float64)
        self.cams = arr(s.cams, C * nc, np.float64)
        self.pts = arr(s.pts, 3 * npt, np.float64)
        self.meas = arr(s.meas, 2 * no, np.float64)
        self.iidx = arr(s.iidx, no, np.int32)
        self.jidx = arr(s.jidx, no, np.int32)
        self.fixmask = arr(s.fixmask, C * nc, np.uint8)
        self.gt_q0 = arr(s.gt_q0, 4 * nc, np.float64)
        self.gt_cams = arr(s.gt_cams, C * nc, np.float64)
        self.gt_pts = arr(s.gt_pts, 3 * npt, np.float64)
        self.sum_t2, self.n_outliers, self.zipf_s = s.sum_t2, s.n_outliers, s.zipf_s
        lib().psba_scene_free(p)


class Context:
    """Device state of one problem (the HIP-backed PSBA_struct)."""

    def __init__(self, cnp, n_cams, n_pts, n_obs, precision=PREC_FP64, device=0):
        L = lib()
        self._h = ct.c_void_p()
        d = Desc(cnp, 3, 2, n_cams, n_pts, n_obs, precision, device)
        _check(L.psba_create(ct.byref(d), ct.byref(self._h)), "psba_create")
        self.cnp, self.n_cams, self.n_pts, self.n_obs = cnp, n_cams, n_pts, n_obs
        self.itno = 0

    @classmethod
    def from_scene(cls, sc, fix=True, **kw):
        c = cls(sc.cnp, sc.n_cams, sc.n_pts, sc.n_obs, **kw)
        c.upload(sc.K, sc.meas, sc.q0, sc.cams, sc.pts)
        c.set_index(sc.iidx, sc.jidx)
        if fix:
            c.set_fixed(sc.fixmask)
        return c

    def close(self):
        if self._h:
            lib().psba_destroy(self._h)
            self._h = ct.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def upload(self, K, impts, initrot, cams, pts):
        self._keep = [np.ascontiguousarray(x, dtype=np.float64) for x in (K, impts, initrot, cams, pts)]
        _check(lib().psba_upload(self._h, *[_dptr(x) for x in self._keep]), "psba_upload")

    def set_index(self, iidx, jidx):
        i = np.ascontiguousarray(iidx, dtype=np.int32)
        j = np.ascontiguousarray(jidx, dtype=np.int32)
        _check(lib().psba_set_index(self._h, _iptr(i), _iptr(j)), "psba_set_index")

    def set_loss(self, kind, scale=1.0):
        """Robust loss on each observation's squared residual norm (LOSS_*)."""
        _check(lib().psba_set_loss(self._h, kind, scale), "psba_set_loss")

    def set_index_device(self, iidx, jidx):
        """The same index built on the device (bit-identical to set_index)."""
        i = np.ascontiguousarray(iidx, dtype=np.int32)
        j = np.ascontiguousarray(jidx, dtype=np.int32)
        _check(lib().psba_set_index_device(self._h, _iptr(i), _iptr(j)), "psba_set_index_device")

    def read_index(self, name):
        which = INDEX[name]
        n = lib().psba_index_size(self._h, which)
        if n < 0:
            raise PsbaError("index not built")
        out = np.empty(max(n, 1), np.int32)
        _check(lib().psba_read_index(self._h, which, _iptr(out), out.size), "psba_read_index")
        return out[:n]

    def set_fixed(self, mask):
        if mask is None:
            _check(lib().psba_set_fixed(self._h, None), "psba_set_fixed")
            return
        m = np.ascontiguousarray(mask, dtype=np.uint8)
        _check(lib().psba_set_fixed(self._h, m.ctypes.data_as(ct.POINTER(ct.c_ubyte))), "psba_set_fixed")

    def read_params(self):
        cams = np.empty(self.cnp * self.n_cams)
        pts = np.empty(3 * self.n_pts)
        _check(lib().psba_read_params(self._h, _dptr(cams), _dptr(pts)), "psba_read_params")
        return cams, pts

    def write_params(self, cams, pts):
        c = np.ascontiguousarray(cams, dtype=np.float64)
        p = np.ascontiguousarray(pts, dtype=np.float64)
        _check(lib().psba_write_params(self._h, _dptr(c), _dptr(p)), "psba_write_params")

    def snapshot(self):
        _check(lib().psba_snapshot_params(self._h), "psba_snapshot_params")

    def restore(self):
        _check(lib().psba_restore_params(self._h), "psba_restore_params")

    def _opts(self, max_iter=50, max_pb_tries=0, flags=0, rel_tol=0.0, refine_steps=0):
        return Opts(max_iter, max_pb_tries, refine_steps, flags, rel_tol)

    def _run(self, fn, trace_cap, opts, extra_init=False):
        tr = (TraceRec * max(trace_cap, 1))()
        n = ct.c_int(0)
        it = ct.c_int(self.itno)
        fe = ct.c_double(0.0)
        if extra_init:
            ie = ct.c_double(0.0)
            rc = fn(self._h, ct.byref(opts), ct.byref(it), ct.byref(ie), ct.byref(fe), tr, trace_cap, ct.byref(n))
        else:
            rc = fn(self._h, ct.byref(opts), ct.byref(it), ct.byref(fe), tr, trace_cap, ct.byref(n))
        _check(rc, fn.__name__)
        self.itno = it.value
        recs = [tr[k] for k in range(min(n.value, trace_cap))]
        out = dict(code=rc, itno=it.value, final_err=fe.value, trace=recs)
        if extra_init:
            out["init_err"] = ie.value
        return out

    def levmar(self, max_iter=50, trace_cap=4096, flags=0, rel_tol=0.0, refine_steps=0):
        """PSBA/levmar.h:6-11 on the device; returns code, itno, final_err, trace.
        refine_steps: FP64 refinement sweeps per step (mixed contexts)."""
        return self._run(lib().psba_levmar, trace_cap, self._opts(max_iter, 0, flags, rel_tol, refine_steps))

    def levmar_entry_cost(self):
        """SSE at the entry of the last levmar call (the reference's initErr)."""
        v = ct.c_double()
        _check(lib().psba_levmar_entry_cost(self._h, ct.byref(v)), "psba_levmar_entry_cost")
        return v.value

    def trust_region(self, max_iter=50, trace_cap=4096, max_pb_tries=0):
        """PSBA/trust_region.h:5-11 on the device."""
        return self._run(lib().psba_trust_region, trace_cap, self._opts(max_iter, max_pb_tries))

    def hybrid(self, max_iter=50, trace_cap=4096):
        """The LM <-> dogleg loop of PSBA/main.cpp:193-208."""
        return self._run(lib().psba_hybrid, trace_cap, self._opts(max_iter), extra_init=True)

    # ---- stages ----
    def hybrid_report(self, max_iter=50):
        """The hybrid loop with the reference's report (PSBA/main.cpp:209-225)
        plus the true RMS, LM/dogleg split and device time per stage."""
        rep = Report()
        it = ct.c_int(self.itno)
        _check(lib().psba_hybrid_report(self._h, ct.byref(self._opts(max_iter)), ct.byref(it), ct.byref(rep)),
               "psba_hybrid_report")
        self.itno = it.value
        out = {k: getattr(rep, k) for k, _ in Report._fields_ if k != "stages"}
        out["stages"] = {k: getattr(rep.stages, k) for k, _ in ReportStages._fields_}
        return out

    def residual(self, trial=False):
        c = ct.c_double()
        _check(lib().psba_stage_residual(self._h, int(trial), ct.byref(c)), "residual")
        return c.value

    def jacobian_normal(self, coeff=1.0, coeff_g=1.0):
        _check(lib().psba_stage_jacobian_normal(self._h, coeff, coeff_g), "jacobian_normal")

    def schur(self, mu):
        _check(lib().psba_stage_schur(self._h, mu), "schur")

    def factor_solve(self):
        ok = ct.c_int()
        _check(lib().psba_stage_factor_solve(self._h, ct.byref(ok)), "factor_solve")
        return bool(ok.value)

    def backsub(self):
        _check(lib().psba_stage_backsub(self._h), "backsub")

    def refine(self, mu, sweeps=1):
        _check(lib().psba_stage_refine(self._h, mu, sweeps), "psba_stage_refine")

    def modchol(self):
        lam = ct.c_double()
        _check(lib().psba_stage_modchol(self._h, ct.byref(lam)), "modchol")
        return lam.value

    def load_S(self, S, b=None):
        """Test hook: S (n x n) as the reduced matrix and b as its right-hand side."""
        Sm = np.ascontiguousarray(S, dtype=np.float64)
        bb = None if b is None else np.ascontiguousarray(b, dtype=np.float64)
        _check(lib().psba_stage_load_S(self._h, _dptr(Sm), None if bb is None else _dptr(bb)), "load_S")

    def vinv_blocks(self, V, mu=0.0):
        """Test hook: the damped 3x3 inverses k_vinv forms for n_pts given
        blocks V (n_pts x 3 x 3); replaces the context's V."""
        Vm = np.ascontiguousarray(V, dtype=np.float64).reshape(self.n_pts, 9)
        out = np.empty((self.n_pts, 9))
        _check(lib().psba_stage_vinv(self._h, _dptr(Vm), float(mu), _dptr(out)), "psba_stage_vinv")
        return out.reshape(self.n_pts, 3, 3)

    def factor_only(self):
        """Factor + solve of the loaded S; returns (ok, x) with x = S^-1 b."""
        ok = self.factor_solve()
        return ok, self.buffer("DP")[: self.cnp * self.n_cams]

    def buffer(self, name):
        w = BUF[name]
        n = lib().psba_buffer_size(self._h, w)
        out = np.empty(n)
        _check(lib().psba_read_buffer(self._h, w, _dptr(out), n), "read_buffer")
        return out

    def info(self):
        i = Info()
        _check(lib().psba_get_info(self._h, ct.byref(i)), "get_info")
        return {k: getattr(i, k) for k, _ in Info._fields_}

    def probe(self, which):
        _check(lib().psba_probe(self._h, which), "probe")

    def probe_read(self):
        t = ct.c_double()
        n = ct.c_int()
        _check(lib().psba_probe_read(self._h, ct.byref(t), ct.byref(n)), "probe_read")
        return t.value, n.value

    def enable_timing(self, on=True):
        _check(lib().psba_enable_timing(self._h, int(on)), "enable_timing")

    def timing(self):
        t = Timing()
        _check(lib().psba_get_timing(self._h, ct.byref(t)), "get_timing")
        return {k: getattr(t, k) for k, _ in Timing._fields_}


def generate_idxs(n_cams, n_pts, n_obs, vmask, with_comm=False):
    """PSBA/misc.cpp:178-218 through the C-ABI; returns (iidx, jidx, blk_idx, cnt, comm|None)."""
    vm = np.ascontiguousarray(vmask, dtype=np.int8).reshape(-1)
    iidx = np.empty(n_obs, np.int32)
    jidx = np.empty(n_obs, np.int32)
    blk = np.empty(n_pts * n_cams, np.int32)
    cnt = np.empty(n_cams * n_cams, np.int32)
    comm = np.full(n_cams * n_cams * n_pts, -7, np.int32) if with_comm else None
    rc = lib().psba_generate_idxs(n_cams, n_pts, n_obs, vm.ctypes.data_as(ct.c_char_p),
                                  _iptr(comm) if comm is not None else None, _iptr(cnt), _iptr(iidx),
                                  _iptr(jidx), _iptr(blk))
    _check(rc, "generate_idxs")
    return iidx, jidx, blk, cnt, comm


def generate_idxs_device(n_cams, n_pts, n_obs, vmask, device=0):
    """generate_idxs on a device; returns (iidx, jidx, blk_idx, cnt)."""
    vm = np.ascontiguousarray(vmask, dtype=np.int8).reshape(-1)
    iidx = np.empty(n_obs, np.int32)
    jidx = np.empty(n_obs, np.int32)
    blk = np.empty(n_pts * n_cams, np.int32)
    cnt = np.empty(n_cams * n_cams, np.int32)
    rc = lib().psba_generate_idxs_device(device, n_cams, n_pts, n_obs, vm.ctypes.data_as(ct.c_char_p), _iptr(cnt),
                                         _iptr(iidx), _iptr(jidx), _iptr(blk))
    _check(rc, "generate_idxs_device")
    return iidx, jidx, blk, cnt


def problem_bytes(cnp, n_cams, n_pts, n_obs, pair_items, precision=PREC_FP64):
    """Device bytes one problem takes inside a batch (an upper bound)."""
    return lib().psba_problem_bytes(cnp, precision, n_cams, n_pts, n_obs, int(pair_items))


def pair_items(iidx):
    """sum over points of t(t+1)/2, t = track length (iidx point-major)."""
    t = np.bincount(np.asarray(iidx, dtype=np.int64))
    return int((t * (t + 1) // 2).sum())


class Batch:
    """Many independent problems solved together (include/psba_batch.h)."""

    def __init__(self, cnp, precision=PREC_FP64, device=0, max_problems=0):
        L = lib()
        self._h = ct.c_void_p()
        d = BatchDesc(cnp, precision, device, max_problems)
        _check(L.psba_batch_create(ct.byref(d), ct.byref(self._h)), "psba_batch_create")
        self.cnp = cnp
        self.sizes = []

    def close(self):
        if self._h:
            lib().psba_batch_destroy(self._h)
            self._h = ct.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def add(self, K, impts, initrot, cams, pts, iidx, jidx, fixmask=None):
        a = [np.ascontiguousarray(x, dtype=np.float64) for x in (K, impts, initrot, cams, pts)]
        i = np.ascontiguousarray(iidx, dtype=np.int32)
        j = np.ascontiguousarray(jidx, dtype=np.int32)
        nc, npt, no = a[3].size // self.cnp, a[4].size // 3, i.size
        fm = None
        if fixmask is not None:
            fm = np.ascontiguousarray(fixmask, dtype=np.uint8)
        slot = ct.c_int(-1)
        _check(lib().psba_batch_add(self._h, nc, npt, no, *[_dptr(x) for x in a], _iptr(i), _iptr(j),
                                    None if fm is None else fm.ctypes.data_as(ct.POINTER(ct.c_ubyte)),
                                    ct.byref(slot)), "psba_batch_add")
        self.sizes.append((nc, npt, no))
        return slot.value

    def add_scene(self, sc, fix=True):
        return self.add(sc.K, sc.meas, sc.q0, sc.cams, sc.pts, sc.iidx, sc.jidx, sc.fixmask if fix else None)

    def reset(self):
        _check(lib().psba_batch_reset(self._h), "psba_batch_reset")
        self.sizes = []

    def commit(self):
        _check(lib().psba_batch_commit(self._h), "psba_batch_commit")

    def nbytes(self):
        return lib().psba_batch_bytes(self._h)

    @staticmethod
    def _opts(max_iter, rel_tol):
        return Opts(max_iter, 0, 0, 0, rel_tol)

    def launch(self, max_iter=100, rel_tol=1e-8):
        _check(lib().psba_batch_launch(self._h, ct.byref(self._opts(max_iter, rel_tol))), "psba_batch_launch")

    def poll(self):
        return _check(lib().psba_batch_poll(self._h), "psba_batch_poll") == 1

    def solve(self, max_iter=100, rel_tol=1e-8):
        r = ct.c_int(0)
        _check(lib().psba_batch_solve(self._h, ct.byref(self._opts(max_iter, rel_tol)), ct.byref(r)),
               "psba_batch_solve")
        return r.value

    def result(self, slot, params=True):
        if not 0 <= slot < len(self.sizes):
            params = False  # the library reports the bad slot
        nc, npt, _ = self.sizes[slot] if params else (0, 0, 0)
        code, it = ct.c_int(0), ct.c_int(0)
        ie, fe = ct.c_double(0), ct.c_double(0)
        cams = np.empty(self.cnp * nc) if params else None
        pts = np.empty(3 * npt) if params else None
        _check(lib().psba_batch_result(self._h, slot, ct.byref(code), ct.byref(it), ct.byref(ie), ct.byref(fe),
                                       _dptr(cams) if params else None, _dptr(pts) if params else None),
               "psba_batch_result")
        out = dict(code=code.value, itno=it.value, init_err=ie.value, final_err=fe.value)
        if params:
            out["cams"], out["pts"] = cams, pts
        return out

    def elapsed(self):
        ms, r = ct.c_double(0), ct.c_int(0)
        _check(lib().psba_batch_elapsed(self._h, ct.byref(ms), ct.byref(r)), "psba_batch_elapsed")
        return ms.value, r.value

    def autotune(self):
        ob, pb = ct.c_int(0), ct.c_int(0)
        _check(lib().psba_batch_autotune(self._h, ct.byref(ob), ct.byref(pb)), "psba_batch_autotune")
        return ob.value, pb.value


def server_clock():
    """Seconds on the steady clock the server stamps requests with."""
    return lib().psba_server_clock()


class Server:
    """Request loop over batches (include/psba_server.h)."""

    def __init__(self, cnp, precision=PREC_FP64, device=0, lanes=0, max_batch=0, hbm_budget=0, max_iter=100,
                 rel_tol=1e-8):
        L = lib()
        self._h = ct.c_void_p()
        d = ServerDesc(cnp, precision, device, lanes, max_batch, int(hbm_budget), max_iter, rel_tol)
        _check(L.psba_server_create(ct.byref(d), ct.byref(self._h)), "psba_server_create")
        self.cnp = cnp
        self._sizes = {}

    def close(self):
        if self._h:
            lib().psba_server_destroy(self._h)
            self._h = ct.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def submit(self, K, impts, initrot, cams, pts, iidx, jidx, fixmask=None):
        a = [np.ascontiguousarray(x, dtype=np.float64) for x in (K, impts, initrot, cams, pts)]
        i = np.ascontiguousarray(iidx, dtype=np.int32)
        j = np.ascontiguousarray(jidx, dtype=np.int32)
        nc, npt, no = a[3].size // self.cnp, a[4].size // 3, i.size
        fm = None if fixmask is None else np.ascontiguousarray(fixmask, dtype=np.uint8)
        rid = ct.c_longlong(0)
        _check(lib().psba_server_submit(self._h, nc, npt, no, *[_dptr(x) for x in a], _iptr(i), _iptr(j),
                                        None if fm is None else fm.ctypes.data_as(ct.POINTER(ct.c_ubyte)),
                                        ct.byref(rid)), "psba_server_submit")
        self._sizes[rid.value] = (nc, npt)
        return rid.value

    def submit_scene(self, sc, fix=True):
        return self.submit(sc.K, sc.meas, sc.q0, sc.cams, sc.pts, sc.iidx, sc.jidx, sc.fixmask if fix else None)

    def step(self):
        return _check(lib().psba_server_step(self._h), "psba_server_step")

    def drain(self):
        _check(lib().psba_server_drain(self._h), "psba_server_drain")

    def result(self, rid, params=True):
        """The request's result (and release), or None while it runs."""
        nc, npt = self._sizes.get(rid, (0, 0))
        code, it = ct.c_int(0), ct.c_int(0)
        ie, fe, ts, td = ct.c_double(0), ct.c_double(0), ct.c_double(0), ct.c_double(0)
        cams = np.empty(self.cnp * nc) if params else None
        pts = np.empty(3 * npt) if params else None
        r = _check(lib().psba_server_result(self._h, rid, ct.byref(code), ct.byref(it), ct.byref(ie), ct.byref(fe),
                                            _dptr(cams) if params else None, _dptr(pts) if params else None,
                                            ct.byref(ts), ct.byref(td)), "psba_server_result")
        if r == 0:
            return None
        self._sizes.pop(rid, None)
        out = dict(code=code.value, itno=it.value, init_err=ie.value, final_err=fe.value, t_submit=ts.value,
                   t_done=td.value)
        if params:
            out["cams"], out["pts"] = cams, pts
        return out

    def completed(self, cap=4096):
        """Ids completed since the last call, in completion order."""
        buf = (ct.c_longlong * cap)()
        n = _check(lib().psba_server_completed(self._h, buf, cap), "psba_server_completed")
        return [buf[k] for k in range(n)]

    def stats(self):
        st = ServerStats()
        _check(lib().psba_server_get_stats(self._h, ct.byref(st)), "psba_server_get_stats")
        return {k: getattr(st, k) for k, _ in ServerStats._fields_}
