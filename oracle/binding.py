"""ctypes view of the CPU oracle (oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product (psba_amd/).  The method
names mirror psba_amd.Context so parity tests read stage by stage.
"""
import ctypes as ct
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")

LOSS_NONE, LOSS_HUBER, LOSS_CAUCHY = 0, 1, 2

SOLVE_CHOLESKY = 0
SOLVE_INVERSE = 1

BUF = dict(EX=0, JA=1, JB=2, U=3, V=4, W=5, G=6, VINV=7, Y=8, S=9, EA=10, DP=11, EB=12, L=13,
           NEWCAMS=14, NEWPTS=15)


class Problem(ct.Structure):
    _fields_ = [("cnp", ct.c_int), ("n_cams", ct.c_int), ("n_pts", ct.c_int), ("n_obs", ct.c_int),
                ("K", ct.POINTER(ct.c_double)), ("q0", ct.POINTER(ct.c_double)),
                ("meas", ct.POINTER(ct.c_double)), ("iidx", ct.POINTER(ct.c_int)),
                ("jidx", ct.POINTER(ct.c_int)), ("fixmask", ct.POINTER(ct.c_ubyte)),
                ("cams", ct.POINTER(ct.c_double)), ("pts", ct.POINTER(ct.c_double))]


FLAG_NO_HANDOFF = 1


class Opts(ct.Structure):
    _fields_ = [("solve_mode", ct.c_int), ("max_iter", ct.c_int), ("max_pb_tries", ct.c_int), ("flags", ct.c_int),
                ("rel_tol", ct.c_double)]


class TraceRec(ct.Structure):
    _fields_ = [("itno", ct.c_int), ("kind", ct.c_int), ("factor_ok", ct.c_int), ("accepted", ct.c_int),
                ("cost", ct.c_double), ("damp", ct.c_double), ("rho", ct.c_double), ("lambda_", ct.c_double),
                ("dp_sq", ct.c_double)]


_lib = None


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("liboracle.so not built; run `python -m psba_amd.build --oracle`")
    L = ct.CDLL(LIB_PATH)
    vp, dp, ip = ct.c_void_p, ct.POINTER(ct.c_double), ct.POINTER(ct.c_int)
    sig = {
        "orc_cnp_supported": (ct.c_int, [ct.c_int]),
        "orc_project": (ct.c_int, [ct.c_int, dp, dp, dp, dp, dp, dp, dp]),
        "orc_generate_idxs": (None, [ct.c_int, ct.c_int, ct.c_char_p, ip, ip, ip, ip, ip]),
        "orc_create": (vp, [ct.POINTER(Problem)]),
        "orc_destroy": (None, [vp]),
        "orc_set_threads": (None, [ct.c_int]),
        "orc_set_sum_order": (None, [ct.c_int]),
        "orc_get_threads": (ct.c_int, []),
        "orc_get_params": (None, [vp, dp, dp]),
        "orc_set_params": (None, [vp, dp, dp]),
        "orc_stage_residual": (ct.c_double, [vp, ct.c_int, dp]),
        "orc_stage_jacobian": (None, [vp]),
        "orc_stage_normal": (None, [vp, ct.c_double, ct.c_double]),
        "orc_max_uvdiag": (ct.c_double, [vp]),
        "orc_stage_vinv": (ct.c_int, [vp, ct.c_double]),
        "orc_stage_schur": (None, [vp, ct.c_double]),
        "orc_stage_ea": (None, [vp]),
        "orc_stage_factor_solve": (ct.c_int, [vp, ct.c_int]),
        "orc_stage_backsub": (None, [vp]),
        "orc_stage_newp": (None, [vp, dp]),
        "orc_stage_commit": (None, [vp]),
        "orc_jmul": (None, [vp, dp, dp]),
        "orc_modchol_lambda": (ct.c_double, [vp, dp]),
        "orc_buffer_size": (ct.c_long, [vp, ct.c_int]),
        "orc_read_buffer": (ct.c_int, [vp, ct.c_int, dp]),
        "orc_levmar": (ct.c_int, [vp, ct.POINTER(Opts), ip, dp, ct.POINTER(TraceRec), ct.c_int, ip]),
        "orc_trust_region": (ct.c_int, [vp, ct.POINTER(Opts), ip, dp, ct.POINTER(TraceRec), ct.c_int, ip]),
        "orc_hybrid": (ct.c_int, [vp, ct.POINTER(Opts), ip, dp, dp, ct.POINTER(TraceRec), ct.c_int, ip]),
        "orc_dense_factor_solve": (ct.c_int, [ct.c_int, ct.c_int, dp, dp, dp, dp]),
        "orc_dense_modchol": (ct.c_double, [ct.c_int, dp, dp, dp]),
        "orc_set_loss": (ct.c_int, [vp, ct.c_int, ct.c_double]),
        "orc_loss_eval": (None, [ct.c_int, ct.c_double, ct.c_double, dp, dp]),
        "orc_vinv3": (ct.c_int, [dp, ct.c_double, dp]),
        "orc_dogleg_step": (ct.c_double, [ct.c_int, ct.c_double, ct.c_double, ct.c_double, ct.c_double, dp, dp,
                                          dp, dp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def _d(a):
    return a.ctypes.data_as(ct.POINTER(ct.c_double))


def _i(a):
    return a.ctypes.data_as(ct.POINTER(ct.c_int))


def set_threads(n):
    lib().orc_set_threads(int(n))


def set_sum_order(descending):
    """Every reduction in ascending (False, the default) or descending index order."""
    lib().orc_set_sum_order(int(bool(descending)))


def project(cnp, K5, q0, cam, M, jac=False):
    """uv (2,), and JA (2 x cnp), JB (2 x 3) when jac."""
    a = [np.ascontiguousarray(x, dtype=np.float64) for x in (K5, q0, cam, M)]
    uv = np.empty(2)
    JA = np.empty(2 * cnp) if jac else None
    JB = np.empty(6) if jac else None
    rc = lib().orc_project(cnp, *[_d(x) for x in a], _d(uv), _d(JA) if jac else None, _d(JB) if jac else None)
    if rc != 0:
        raise ValueError("unsupported cnp")
    return (uv, JA.reshape(2, cnp), JB.reshape(2, 3)) if jac else uv


def dense_factor_solve(S, b, solve_mode=SOLVE_CHOLESKY):
    """(ok, x, L) for S x = b (lower triangle of S used)."""
    Sm = np.ascontiguousarray(S, np.float64)
    n = Sm.shape[0]
    bb = np.ascontiguousarray(b, np.float64)
    x = np.zeros(n)
    L = np.zeros((n, n))
    ok = lib().orc_dense_factor_solve(n, solve_mode, _d(Sm), _d(bb), _d(x), _d(L))
    return bool(ok), x, L


def vinv3(V, mu=0.0):
    """(inverse, small-determinant branch taken) of one damped 3x3 block."""
    Vm = np.ascontiguousarray(V, np.float64).reshape(9)
    out = np.empty(9)
    br = lib().orc_vinv3(_d(Vm), float(mu), _d(out))
    return out.reshape(3, 3), bool(br)


def dogleg_step(P_U, P_B, g, pUtBpU, pUtBpB, pBtBpB, delta):
    """(p, reported norm) of the oracle's dogleg / 2-D subspace step."""
    PU = np.ascontiguousarray(P_U, np.float64)
    PB = np.ascontiguousarray(P_B, np.float64)
    gg = np.ascontiguousarray(g, np.float64)
    p = np.empty_like(PU)
    nrm = lib().orc_dogleg_step(PU.size, pUtBpU, pUtBpB, pBtBpB, delta, _d(PU), _d(PB), _d(gg), _d(p))
    return p, nrm


def dense_modchol(S):
    """(lambda, L, E) of the dogleg's modified Cholesky."""
    Sm = np.ascontiguousarray(S, np.float64)
    n = Sm.shape[0]
    L = np.zeros((n, n))
    E = np.zeros(n)
    lam = lib().orc_dense_modchol(n, _d(Sm), _d(L), _d(E))
    return lam, L, E


def generate_idxs(n_cams, n_pts, vmask, with_comm=True):
    vm = np.ascontiguousarray(vmask, dtype=np.int8).reshape(-1)
    n_obs = int(np.count_nonzero(vm))
    iidx = np.empty(n_obs, np.int32)
    jidx = np.empty(n_obs, np.int32)
    blk = np.empty(n_pts * n_cams, np.int32)
    cnt = np.empty(n_cams * n_cams, np.int32)
    comm = np.full(n_cams * n_cams * n_pts, -7, np.int32) if with_comm else None
    lib().orc_generate_idxs(n_cams, n_pts, vm.ctypes.data_as(ct.c_char_p), _i(comm) if with_comm else None,
                            _i(cnt), _i(iidx), _i(jidx), _i(blk))
    return iidx, jidx, blk, cnt, comm


class Oracle:
    def __init__(self, cnp, K, q0, meas, iidx, jidx, cams, pts, fixmask=None):
        self._arr = dict(K=np.ascontiguousarray(K, np.float64), q0=np.ascontiguousarray(q0, np.float64),
                         meas=np.ascontiguousarray(meas, np.float64), iidx=np.ascontiguousarray(iidx, np.int32),
                         jidx=np.ascontiguousarray(jidx, np.int32), cams=np.ascontiguousarray(cams, np.float64),
                         pts=np.ascontiguousarray(pts, np.float64))
        a = self._arr
        n_cams, n_pts, n_obs = len(a["K"]) // 5, len(a["pts"]) // 3, len(a["iidx"])
        fm = None
        if fixmask is not None:
            self._fm = np.ascontiguousarray(fixmask, np.uint8)
            fm = self._fm.ctypes.data_as(ct.POINTER(ct.c_ubyte))
        p = Problem(cnp, n_cams, n_pts, n_obs, _d(a["K"]), _d(a["q0"]), _d(a["meas"]), _i(a["iidx"]),
                    _i(a["jidx"]), fm, _d(a["cams"]), _d(a["pts"]))
        self._h = lib().orc_create(ct.byref(p))
        if not self._h:
            raise ValueError("oracle: bad problem")
        self.cnp, self.n_cams, self.n_pts, self.n_obs = cnp, n_cams, n_pts, n_obs
        self.itno = 0

    @classmethod
    def from_scene(cls, sc, fix=True):
        return cls(sc.cnp, sc.K, sc.q0, sc.meas, sc.iidx, sc.jidx, sc.cams, sc.pts, sc.fixmask if fix else None)

    def close(self):
        if self._h:
            lib().orc_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def read_params(self):
        c = np.empty(self.cnp * self.n_cams)
        p = np.empty(3 * self.n_pts)
        lib().orc_get_params(self._h, _d(c), _d(p))
        return c, p

    def write_params(self, cams, pts):
        c = np.ascontiguousarray(cams, np.float64)
        p = np.ascontiguousarray(pts, np.float64)
        lib().orc_set_params(self._h, _d(c), _d(p))

    def set_loss(self, kind, scale=1.0):
        if lib().orc_set_loss(self._h, kind, scale) != 0:
            raise ValueError("bad loss")

    def _run(self, fn, trace_cap, solve_mode, max_iter, max_pb_tries=0, init=False, flags=0, rel_tol=0.0):
        o = Opts(solve_mode, max_iter, max_pb_tries, flags, rel_tol)
        tr = (TraceRec * max(trace_cap, 1))()
        n = ct.c_int(0)
        it = ct.c_int(self.itno)
        fe = ct.c_double(0)
        if init:
            ie = ct.c_double(0)
            rc = fn(self._h, ct.byref(o), ct.byref(it), ct.byref(ie), ct.byref(fe), tr, trace_cap, ct.byref(n))
        else:
            rc = fn(self._h, ct.byref(o), ct.byref(it), ct.byref(fe), tr, trace_cap, ct.byref(n))
        self.itno = it.value
        out = dict(code=rc, itno=it.value, final_err=fe.value, trace=[tr[k] for k in range(min(n.value, trace_cap))])
        if init:
            out["init_err"] = ie.value
        return out

    def levmar(self, max_iter=50, trace_cap=4096, solve_mode=SOLVE_CHOLESKY, flags=0, rel_tol=0.0):
        return self._run(lib().orc_levmar, trace_cap, solve_mode, max_iter, flags=flags, rel_tol=rel_tol)

    def trust_region(self, max_iter=50, trace_cap=4096, solve_mode=SOLVE_CHOLESKY, max_pb_tries=0):
        return self._run(lib().orc_trust_region, trace_cap, solve_mode, max_iter, max_pb_tries)

    def hybrid(self, max_iter=50, trace_cap=4096, solve_mode=SOLVE_CHOLESKY):
        return self._run(lib().orc_hybrid, trace_cap, solve_mode, max_iter, init=True)

    # stages
    def residual(self, trial=False):
        return lib().orc_stage_residual(self._h, int(trial), None)

    def jacobian_normal(self, coeff=1.0, coeff_g=1.0):
        lib().orc_stage_jacobian(self._h)
        lib().orc_stage_normal(self._h, coeff, coeff_g)

    def max_uvdiag(self):
        return lib().orc_max_uvdiag(self._h)

    def schur(self, mu):
        lib().orc_stage_vinv(self._h, mu)
        lib().orc_stage_schur(self._h, mu)
        lib().orc_stage_ea(self._h)

    def factor_solve(self, solve_mode=SOLVE_CHOLESKY):
        return bool(lib().orc_stage_factor_solve(self._h, solve_mode))

    def backsub(self):
        lib().orc_stage_backsub(self._h)

    def newp(self, dp):
        d = np.ascontiguousarray(dp, np.float64)
        lib().orc_stage_newp(self._h, _d(d))

    def commit(self):
        lib().orc_stage_commit(self._h)

    def jmul(self, x):
        xx = np.ascontiguousarray(x, np.float64)
        out = np.empty(2 * self.n_obs)
        lib().orc_jmul(self._h, _d(xx), _d(out))
        return out

    def modchol(self):
        E = np.empty(self.cnp * self.n_cams)
        lam = lib().orc_modchol_lambda(self._h, _d(E))
        return lam, E

    def buffer(self, name):
        w = BUF[name]
        n = lib().orc_buffer_size(self._h, w)
        out = np.empty(n)
        lib().orc_read_buffer(self._h, w, _d(out))
        return out


def loss_eval(kind, scale, s):
    """(rho(s), rho'(s)) of the oracle's robust loss."""
    r, d = ct.c_double(), ct.c_double()
    lib().orc_loss_eval(kind, scale, s, ct.byref(r), ct.byref(d))
    return r.value, d.value
