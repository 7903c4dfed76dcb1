"""ctypes binding of the CPU oracle (oracle/_build/liboracle.so).

Test infrastructure: the oracle is the checker the HIP path is compared with.
Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "oracle", "_build", "liboracle.so")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


class Layout(C.Structure):
    _fields_ = [
        ("kappa", C.c_double), ("eps", C.c_double), ("eps1", C.c_double),
        ("qorder", C.c_int), ("n", C.c_int),
        ("vtx", _dp), ("nrm", _dp), ("area", _dp),
        ("depth", C.c_int), ("height", C.c_int), ("maxOrder", C.c_int),
        ("ordMom", _ip), ("ordM2L", _ip),
        ("nCube", C.c_int), ("levStart", _ip), ("levCount", _ip),
        ("cubeLevel", _ip), ("cubeX", _dp), ("kidStart", _ip), ("nKids", _ip),
        ("nLeaf", C.c_int), ("leafStart", _ip), ("leafCount", _ip),
        ("nearOff", _ip), ("nearSrc", _ip),
        ("nM2L", C.c_int), ("m2lSrc", _ip), ("m2lDst", _ip),
        ("nThreads", C.c_int),
    ]


class CTree(C.Structure):
    _fields_ = [
        ("nc", C.c_int), ("depth", C.c_int),
        ("levStart", _ip), ("levCount", _ip), ("level", _ip),
        ("x", _dp), ("eRad", _dp), ("kidStart", _ip), ("nKids", _ip),
        ("chgStart", _ip), ("nChgs", _ip), ("chgIdx", _ip),
        ("nChar", C.c_int), ("pos", _dp), ("chr", _dp), ("theta", C.c_double),
    ]


def build():
    subprocess.check_call(["make", "-s", "-f", "oracle/Makefile"], cwd=ROOT)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        L.fbo_create.restype = C.c_void_p
        L.fbo_create.argtypes = [C.POINTER(Layout)]
        L.fbo_destroy.argtypes = [C.c_void_p]
        L.fbo_q2m.restype = _dp
        L.fbo_q2m.argtypes = [C.c_void_p, C.c_int, C.c_int]
        L.fbo_apply.argtypes = [C.c_void_p, C.c_double, _dp, C.c_double, _dp, C.c_int]
        L.fbo_matvec.argtypes = [C.c_void_p, C.c_double, _dp, C.c_double, _dp]
        L.fbo_psolve_diag.argtypes = [C.c_void_p, _dp, _dp]
        L.fbo_psolve.argtypes = [C.c_void_p, _dp, _dp]
        L.fbo_set_precond.restype = C.c_int
        L.fbo_set_precond.argtypes = [C.c_void_p, C.c_int]
        L.fbo_rhs_tree.argtypes = [C.c_void_p, C.c_void_p, C.c_double, _dp]
        L.fbo_set_direct.restype = C.c_int
        L.fbo_set_direct.argtypes = [C.c_void_p, C.c_int]
        L.fbo_expansions.argtypes = [C.c_void_p, _dp, _dp]
        L.fbo_gmres.restype = C.c_int
        L.fbo_gmres.argtypes = [C.c_void_p, _dp, _dp, C.c_int, C.c_double, C.c_int,
                                _ip, _dp, _dp]
        L.fbo_rhs_direct.argtypes = [C.c_void_p, C.c_int, _dp, _dp, _dp]
        L.fbo_energy_tree.restype = C.c_double
        L.fbo_energy_tree.argtypes = [C.c_void_p, C.POINTER(CTree), _dp]
        L.fbo_ctree_moments.argtypes = [C.c_void_p, C.POINTER(CTree), _dp]
        L.fbo_ker4.argtypes = [C.c_double, C.c_double, _dp, _dp, _dp, _dp]
        L.fbo_panel_ia0.argtypes = [C.POINTER(Layout), C.c_int, C.c_int, _dp]
        L.fbo_derivs.argtypes = [C.c_double, C.c_int, _dp, _dp, _dp]
        L.fbo_common_vtx.restype = C.c_int
        L.fbo_common_vtx.argtypes = [_dp, _dp, _ip, _ip]
        L.fbo_gl_table.argtypes = [C.c_int, _dp, _dp]
        L.fbo_nmom.restype = C.c_int
        L.fbo_nmom.argtypes = [C.c_int]
        _lib = L
    return _lib


def dptr(a):
    return a.ctypes.data_as(_dp)


def iptr(a):
    return a.ctypes.data_as(_ip)


def nmom(p):
    return (p + 1) * (p + 2) * (p + 3) // 6


class Oracle:
    """Owns numpy copies of a layout (dict of arrays) and an oracle context.

    layout keys: kappa eps eps1 qorder vtx(n,9) nrm(n,3) area(n) depth height
    maxOrder ordMom ordM2L levStart levCount cubeLevel cubeX(nc,3) kidStart nKids
    leafStart leafCount nearOff nearSrc m2lSrc m2lDst
    """

    def __init__(self, lay, threads=8):
        self.keep = {}

        def d(k):
            a = np.ascontiguousarray(lay[k], dtype=np.float64)
            self.keep[k] = a
            return dptr(a)

        def i(k):
            a = np.ascontiguousarray(lay[k], dtype=np.int32)
            self.keep[k] = a
            return iptr(a)

        self.n = int(len(lay["area"]))
        self.lay = Layout(
            kappa=lay["kappa"], eps=lay["eps"], eps1=lay["eps1"], qorder=lay["qorder"], n=self.n,
            vtx=d("vtx"), nrm=d("nrm"), area=d("area"),
            depth=lay["depth"], height=lay["height"], maxOrder=lay["maxOrder"],
            ordMom=i("ordMom"), ordM2L=i("ordM2L"),
            nCube=int(len(lay["cubeLevel"])), levStart=i("levStart"), levCount=i("levCount"),
            cubeLevel=i("cubeLevel"), cubeX=d("cubeX"), kidStart=i("kidStart"), nKids=i("nKids"),
            nLeaf=int(len(lay["leafStart"])), leafStart=i("leafStart"), leafCount=i("leafCount"),
            nearOff=i("nearOff"), nearSrc=i("nearSrc"),
            nM2L=int(len(lay["m2lSrc"])), m2lSrc=i("m2lSrc"), m2lDst=i("m2lDst"),
            nThreads=threads)
        self.ctx = lib().fbo_create(C.byref(self.lay))
        if not self.ctx:
            raise RuntimeError("fbo_create failed")
        self.maxOrder = lay["maxOrder"]

    def __del__(self):
        if getattr(self, "ctx", None):
            lib().fbo_destroy(self.ctx)
            self.ctx = None

    def apply(self, x, alpha=1.0, beta=0.0, y=None, parts=3):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.zeros(2 * self.n) if y is None else np.array(y, dtype=np.float64)
        lib().fbo_apply(self.ctx, alpha, dptr(x), beta, dptr(y), parts)
        return y

    def matvec(self, x, alpha=1.0, beta=0.0, y=None):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.zeros(2 * self.n) if y is None else np.array(y, dtype=np.float64)
        lib().fbo_matvec(self.ctx, alpha, dptr(x), beta, dptr(y))
        return y

    def set_precond(self, mode):
        """3: diagonal; 0, 1, 2: the leaf-parent block LU (src/precond_fmm.c)."""
        rc = lib().fbo_set_precond(self.ctx, mode)
        if rc != 0:
            raise RuntimeError(f"fbo_set_precond({mode}) = {rc}")

    def set_direct(self, on):
        """-r=1/2: matvec and GMRES on the dense direct operator instead of the FMM."""
        if lib().fbo_set_direct(self.ctx, int(on)) != 0:
            raise RuntimeError("fbo_set_direct failed")

    def psolve_any(self, x):
        """The preconditioner chosen by set_precond."""
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.empty_like(x)
        lib().fbo_psolve(self.ctx, dptr(y), dptr(x))
        return y

    def psolve(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.zeros_like(x)
        lib().fbo_psolve_diag(self.ctx, dptr(y), dptr(x))
        return y

    def expansions(self, ncube):
        nm = nmom(self.maxOrder)
        mom = np.zeros(ncube * 2 * nm)
        lec = np.zeros(ncube * 4 * nm)
        lib().fbo_expansions(self.ctx, dptr(mom), dptr(lec))
        return mom, lec

    def q2m(self, leaf, job, count, p):
        ptr = lib().fbo_q2m(self.ctx, leaf, job)
        return np.ctypeslib.as_array(ptr, shape=(count * nmom(p),)).copy()

    def gmres(self, b, x0, restart=30, tol=1e-6, maxit=100):
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.array(x0, dtype=np.float64)
        it = C.c_int(0)
        res = C.c_double(0)
        hist = np.zeros(maxit * 2 + 2)
        info = lib().fbo_gmres(self.ctx, dptr(b), dptr(x), restart, tol, maxit,
                               C.byref(it), C.byref(res), dptr(hist))
        return x, info, it.value, res.value, hist

    def rhs(self, pos, chr_):
        pos = np.ascontiguousarray(pos, dtype=np.float64)
        chr_ = np.ascontiguousarray(chr_, dtype=np.float64)
        b = np.zeros(2 * self.n)
        lib().fbo_rhs_direct(self.ctx, len(chr_), dptr(pos), dptr(chr_), dptr(b))
        return b

    def panel_ia0(self, x, y):
        out = np.zeros(4)
        lib().fbo_panel_ia0(C.byref(self.lay), x, y, dptr(out))
        return out

    def _ctree(self, ct, pos, chr_, theta=0.8):
        keep = {}

        def i(k):
            a = np.ascontiguousarray(ct[k], dtype=np.int32)
            keep[k] = a
            return iptr(a)

        def d(k, v):
            a = np.ascontiguousarray(v, dtype=np.float64)
            keep[k] = a
            return dptr(a)

        T = CTree(nc=int(len(ct["level"])), depth=int(ct["depth"]),
                  levStart=i("levStart"), levCount=i("levCount"), level=i("level"),
                  x=d("x", ct["x"]), eRad=d("eRad", ct["eRad"]),
                  kidStart=i("kidStart"), nKids=i("nKids"),
                  chgStart=i("chgStart"), nChgs=i("nChgs"), chgIdx=i("chgIdx"),
                  nChar=int(len(chr_)), pos=d("pos", pos), chr=d("chr", chr_), theta=theta)
        return T, keep

    def energy(self, ct, pos, chr_, sigma, theta=0.8):
        T, keep = self._ctree(ct, pos, chr_, theta)
        s = np.ascontiguousarray(sigma, dtype=np.float64)
        return lib().fbo_energy_tree(self.ctx, C.byref(T), dptr(s))

    def rhs_tree(self, ct, pos, chr_, theta=0.8):
        T, keep = self._ctree(ct, pos, chr_, theta)
        b = np.zeros(2 * self.n)
        lib().fbo_rhs_tree(self.ctx, C.byref(T), theta, dptr(b))
        return b

    def ctree_moments(self, ct, pos, chr_):
        T, keep = self._ctree(ct, pos, chr_)
        mom = np.zeros(T.nc * nmom(self.maxOrder))
        lib().fbo_ctree_moments(self.ctx, C.byref(T), dptr(mom))
        return mom


def ker4(kappa, eps, r, nx, ny):
    out = np.zeros(4)
    r, nx, ny = (np.ascontiguousarray(a, dtype=np.float64) for a in (r, nx, ny))
    lib().fbo_ker4(kappa, eps, dptr(r), dptr(nx), dptr(ny), dptr(out))
    return out


def derivs(kappa, order, r):
    nm = nmom(order)
    g0 = np.zeros(nm)
    gk = np.zeros(nm)
    r = np.ascontiguousarray(r, dtype=np.float64)
    lib().fbo_derivs(kappa, order, dptr(r), dptr(g0), dptr(gk))
    return g0, gk


def gl_table(q):
    t = np.zeros(q)
    w = np.zeros(q)
    lib().fbo_gl_table(q, dptr(t), dptr(w))
    return t, w
