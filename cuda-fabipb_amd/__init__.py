"""fabipb_amd - Python host mirror of the fabipb MI355X build.

The product is C and HIP: lib/libfabipb_host.so (mesh input, octree,
interaction lists, charge tree) and lib/libfabipb_hip.so (the device operator
behind the flat C ABI of include/fabipb_hip.h).  This module binds both with
ctypes and strings them together the way the reference driver
src/fabipb.c:804-1229 does: load panels -> tree -> operator -> RHS -> GMRES ->
panel-tree energy.  PyTorch only supplies device memory and the stream.

The HIP library must be built (make -C cuda-fabipb_amd); nothing here falls
back to a CPU path.
"""
import ctypes as C
import math
import os
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBDIR = os.path.join(HERE, "lib")
BINDIR = os.path.join(HERE, "bin")
ENERGY_SCALE = 2.0 * math.pi * 332.0716   # src/fabipb.c:810, 1207

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_vp = C.c_void_p


class FabipbError(RuntimeError):
    pass


_host = None
_hip = None


def _load(name):
    path = os.path.join(LIBDIR, name)
    if not os.path.exists(path):
        raise FabipbError(f"{path} is not built; run `make -C cuda-fabipb_amd`")
    return C.CDLL(path)


def host_lib():
    global _host
    if _host is None:
        L = _load("libfabipb_host.so")
        L.fb_mesh_load.restype = _vp
        L.fb_mesh_load.argtypes = [C.c_char_p, C.c_int]
        L.fb_mesh_free.argtypes = [_vp]
        L.fb_host_error.restype = C.c_char_p
        L.fb_tree_build.restype = _vp
        L.fb_tree_build.argtypes = [_vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double]
        L.fb_tree_build_leaves.restype = _vp
        L.fb_tree_build_leaves.argtypes = [_vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, _ip, _dp, _dp,
                                           _dp, _ip, C.c_int, _vp, _ip, _dp]
        L.fb_tree_free.argtypes = [_vp]
        L.fb_ctree_build.restype = _vp
        L.fb_ctree_build.argtypes = [_vp, C.c_int]
        L.fb_ctree_free.argtypes = [_vp]
        for f in ("fb_tree_array", "fb_ctree_array", "fb_mesh_array"):
            getattr(L, f).restype = _vp
            getattr(L, f).argtypes = [_vp, C.c_char_p, _ip, _ip]
        L.fb_charge_tree_order.restype = C.c_int
        L.fb_charge_tree_order.argtypes = [_vp, C.c_int]
        L.fb_should_use_tree_rhs.restype = C.c_int
        L.fb_should_use_tree_rhs.argtypes = [C.c_int, C.c_int, C.c_int]
        _host = L
    return _host


def hip_lib():
    global _hip
    if _hip is None:
        L = _load("libfabipb_hip.so")
        L.fbh_available.restype = C.c_int
        L.fbh_last_error.restype = C.c_char_p
        L.fbh_tree_panels.restype = C.c_int
        L.fbh_tree_panels.argtypes = [C.c_int, _dp, _dp, _dp, _ip, C.c_int, _ip, _dp, _dp, _dp, _ip, _ip, _vp,
                                      _ip, _dp]
        L.fbh_create.restype = _vp
        L.fbh_create.argtypes = [C.c_double, C.c_double, C.c_double, C.c_int,
                                 C.c_int, _dp, _dp, _dp,
                                 C.c_int, C.c_int, _ip, _ip,
                                 C.c_int, _ip, _ip, _ip, _dp, _ip, _ip, _ip,
                                 C.c_int, _ip, _ip, _ip, _ip, _ip, _ip, _vp]
        L.fbh_destroy.argtypes = [_vp]
        L.fbh_matvec_dev.restype = C.c_int
        L.fbh_matvec_dev.argtypes = [_vp, C.c_double, _vp, C.c_double, _vp]
        L.fbh_matvec.restype = C.c_int
        L.fbh_matvec.argtypes = [_vp, C.c_double, _dp, C.c_double, _dp]
        L.fbh_near_apply.restype = C.c_int
        L.fbh_near_apply.argtypes = [_vp, C.c_double, _dp, _dp]
        L.fbh_far_apply.restype = C.c_int
        L.fbh_far_apply.argtypes = [_vp, C.c_double, _dp, C.c_double, _dp]
        L.fbh_expansions.restype = C.c_int
        L.fbh_expansions.argtypes = [_vp, _dp, _dp]
        L.fbh_q2m_matrices.restype = C.c_int
        L.fbh_q2m_matrices.argtypes = [_vp, _dp, _dp]
        L.fbh_near_coeff.restype = C.c_int
        L.fbh_near_coeff.argtypes = [_vp, C.c_int, C.c_int, _dp]
        L.fbh_psolve_diag_dev.restype = C.c_int
        L.fbh_psolve_diag_dev.argtypes = [_vp, _vp, _vp]
        L.fbh_gmres_dev.restype = C.c_int
        L.fbh_gmres_dev.argtypes = [_vp, _vp, _vp, C.c_int, C.c_double, C.c_int,
                                    _ip, _dp, _dp, _ip]
        L.fbh_set_direct.restype = C.c_int
        L.fbh_set_direct.argtypes = [_vp, C.c_int]
        L.fbh_direct_apply.restype = C.c_int
        L.fbh_direct_apply.argtypes = [_vp, C.c_double, C.c_double, _dp, _dp]
        L.fbh_set_precond.restype = C.c_int
        L.fbh_set_precond.argtypes = [_vp, C.c_int]
        L.fbh_psolve_dev.restype = C.c_int
        L.fbh_psolve_dev.argtypes = [_vp, _vp, _vp]
        L.fbh_enable_fp32.restype = C.c_int
        L.fbh_enable_fp32.argtypes = [_vp]
        L.fbh_matvec32_dev.restype = C.c_int
        L.fbh_matvec32_dev.argtypes = [_vp, C.c_double, _vp, C.c_double, _vp]
        L.fbh_gmres_refined_dev.restype = C.c_int
        L.fbh_gmres_refined_dev.argtypes = [_vp, _vp, _vp, C.c_int, C.c_double, C.c_int, C.c_int,
                                            _ip, _dp, _dp, _ip, _ip, _ip]
        L.fbh_gmres_refined_energy_dev.restype = C.c_int
        L.fbh_gmres_refined_energy_dev.argtypes = [_vp, _vp, _vp, C.c_int, C.c_double, C.c_int, C.c_int,
                                                   C.c_double, _ip, _dp, _dp, _ip, _ip, _ip, _dp]
        L.fbh_rhs.restype = C.c_int
        L.fbh_rhs.argtypes = [_vp, C.c_int, _dp, _dp, _dp]
        L.fbh_rhs_dev.restype = C.c_int
        L.fbh_rhs_dev.argtypes = [_vp, C.c_int, _dp, _dp, _vp]
        L.fbh_rhs_dev_noalloc.restype = C.c_int
        L.fbh_rhs_dev_noalloc.argtypes = [_vp, C.c_int, _dp, _dp, _vp]
        L.fbh_energy_dev.restype = C.c_int
        L.fbh_energy_dev.argtypes = [_vp, C.c_int, _dp, _dp, C.c_int, _ip, _dp, _dp, _ip, _ip,
                                     _ip, _ip, _ip, _ip, C.c_double, _vp, _dp]
        L.fbh_rhs_tree_dev.restype = C.c_int
        L.fbh_rhs_tree_dev.argtypes = [_vp, C.c_int, _dp, _dp, C.c_int, _ip, _dp, _dp, _ip, _ip,
                                       _ip, _ip, _ip, _ip, C.c_double, _vp]
        L.fbh_set_charges.restype = C.c_int
        L.fbh_set_charges.argtypes = [_vp, C.c_int, _dp, _dp, C.c_int, _ip, _dp, _dp, _ip, _ip, _ip, _ip, _ip, _ip]
        L.fbh_rhs_res_dev.restype = C.c_int
        L.fbh_rhs_res_dev.argtypes = [_vp, _vp]
        L.fbh_rhs32_res_dev.restype = C.c_int
        L.fbh_rhs32_res_dev.argtypes = [_vp, _vp]
        L.fbh_rhs_tree_res_dev.restype = C.c_int
        L.fbh_rhs_tree_res_dev.argtypes = [_vp, C.c_double, _vp]
        L.fbh_energy_res_dev.restype = C.c_int
        L.fbh_energy_res_dev.argtypes = [_vp, C.c_double, _vp, _dp]
        L.fbh_energy32_res_dev.restype = C.c_int
        L.fbh_energy32_res_dev.argtypes = [_vp, C.c_double, _vp, _dp]
        L.fbh_matvec_t_dev.restype = C.c_int
        L.fbh_matvec_t_dev.argtypes = [_vp, C.c_double, _vp, C.c_double, _vp]
        L.fbh_gmres_t_dev.restype = C.c_int
        L.fbh_gmres_t_dev.argtypes = [_vp, _vp, _vp, C.c_int, C.c_double, C.c_int, _ip, _dp, _dp, _ip]
        L.fbh_energy_direct_dev.restype = C.c_int
        L.fbh_energy_direct_dev.argtypes = [_vp, C.c_int, _dp, _dp, _vp, _vp, _dp]
        L.fbh_gradient_dev.restype = C.c_int
        L.fbh_gradient_dev.argtypes = [_vp, C.c_int, _dp, _dp, _vp, _vp, C.c_int, C.c_double, C.c_int,
                                       _dp, _dp, _ip, _dp, _ip]
        L.fbh_gradient_tree_dev.restype = C.c_int
        L.fbh_gradient_tree_dev.argtypes = [_vp, C.c_int, _dp, _dp, C.c_int, _ip, _dp, _dp, _ip, _ip, _ip, _ip,
                                            _ip, _ip, C.c_double, C.c_int, _vp, _vp, C.c_int, C.c_int, C.c_double,
                                            C.c_int, _dp, _dp, _ip, _dp, _ip]
        L.fbh_tree_adjoint_dev.restype = C.c_int
        L.fbh_tree_adjoint_dev.argtypes = [_vp, C.c_int, _dp, _dp, C.c_int, _ip, _dp, _dp, _ip, _ip, _ip, _ip,
                                           _ip, _ip, C.c_double, C.c_int, _vp, _dp]
        L.fbh_profile.restype = C.c_int
        L.fbh_profile.argtypes = [_vp, _vp, _vp, C.c_int, _dp]
        L.fbh_profile32.restype = C.c_int
        L.fbh_profile32.argtypes = [_vp, _vp, _vp, C.c_int, _dp]
        L.fbh_near_timing.restype = C.c_int
        L.fbh_near_timing.argtypes = [_vp, C.c_int]
        L.fbh_near_timing_read.restype = C.c_int
        L.fbh_near_timing_read.argtypes = [_vp, C.POINTER(C.c_longlong), _dp]
        L.fbh_stats.restype = C.c_int
        L.fbh_stats.argtypes = [_vp, C.POINTER(C.c_longlong)]
        L.fbh_stream.restype = _vp
        L.fbh_stream.argtypes = [_vp]
        L.fbh_set_rhs_factor.restype = C.c_int
        L.fbh_set_rhs_factor.argtypes = [_vp, C.c_double]
        L.fbh_pair_coeffs.restype = C.c_int
        L.fbh_pair_coeffs.argtypes = [_vp, C.c_longlong, _ip, _ip, _dp, _dp, _dp, _dp]
        L.fbh_panel_pair_coeffs.restype = C.c_int
        L.fbh_panel_pair_coeffs.argtypes = [C.c_int, _dp, _dp, _dp, C.c_double, C.c_double, C.c_int, C.c_longlong,
                                            _ip, _ip, _dp, _dp, _dp, _dp]
        _hip = L
    return _hip


def _hip_err():
    return hip_lib().fbh_last_error().decode()


def _d(a):
    return a.ctypes.data_as(_dp)


def _i(a):
    return a.ctypes.data_as(_ip)


def _ptr(fn, h, name, typ):
    """An array of a host object by pointer (no copy) and its length."""
    cnt, isd = C.c_int(0), C.c_int(0)
    p = fn(h, name.encode(), C.byref(cnt), C.byref(isd))
    return C.cast(p, typ), cnt.value


def _arr(fn, h, name):
    cnt, isd = C.c_int(0), C.c_int(0)
    p = fn(h, name.encode(), C.byref(cnt), C.byref(isd))
    if not p or cnt.value == 0:
        return np.zeros(0, np.float64 if isd.value else np.int32)
    ct = C.c_double if isd.value else C.c_int
    return np.ctypeslib.as_array(C.cast(p, C.POINTER(ct)), shape=(cnt.value,)).copy()


class Mesh:
    """Panels and charges from <base>.vert/.face/.pqr (src/input.c:320-821)."""

    def __init__(self, base, geometric_normals=False):
        L = host_lib()
        self.h = L.fb_mesh_load(base.encode(), int(geometric_normals))
        if not self.h:
            raise FabipbError(L.fb_host_error().decode())
        self.pos = _arr(L.fb_mesh_array, self.h, "pos").reshape(-1, 3)
        self.chr = _arr(L.fb_mesh_array, self.h, "chr")
        self.n = len(_arr(L.fb_mesh_array, self.h, "area"))

    def __del__(self):
        if getattr(self, "h", None):
            host_lib().fb_mesh_free(self.h)
            self.h = None


TREE_KEYS = ["ordMom", "ordM2L", "perm", "vtx", "nrm", "area", "face", "levStart", "levCount",
             "cubeLevel", "kidStart", "nKids", "parent", "cubeX", "eRad", "leafStart", "leafCount",
             "nearOff", "nearSrc", "m2lSrc", "m2lDst", "m2lOff"]


class Tree:
    """Octree, interaction lists and flattened layouts (src/gkSetup.c, src/fmm.c:234-524)."""

    def __init__(self, mesh, depth, order=6, height=2, order_mom=0, sep=0.8, device=False):
        """device: the panel pass (leaf of every panel, leaf order, leaf boxes) on the GPU
        (fbh_tree_panels), the cubes and lists above the leaves on the host; the same arrays."""
        L = host_lib()
        if device:
            H = hip_lib()
            vtx, _ = _ptr(L.fb_mesh_array, mesh.h, "vtx", _dp)
            nrm, _ = _ptr(L.fb_mesh_array, mesh.h, "nrm", _dp)
            area, n = _ptr(L.fb_mesh_array, mesh.h, "area", _dp)
            face, _ = _ptr(L.fb_mesh_array, mesh.h, "face", _ip)
            perm, faceR, cnt = np.empty(n, np.int32), np.empty(n, np.int32), np.empty(n, np.int32)
            vtxR, nrmR, areaR = np.empty(9 * n), np.empty(3 * n), np.empty(n)
            key, box, nleaf = np.empty(n, np.uint32), np.empty(6 * n), C.c_int(0)
            if not H.fbh_tree_panels(n, vtx, nrm, area, face, depth, _i(perm), _d(vtxR), _d(nrmR),
                                     _d(areaR), _i(faceR), C.byref(nleaf), key.ctypes.data_as(_vp), _i(cnt),
                                     _d(box)):
                raise FabipbError(H.fbh_last_error().decode())
            self.h = L.fb_tree_build_leaves(mesh.h, depth, height, order, order_mom, sep, _i(perm), _d(vtxR),
                                            _d(nrmR), _d(areaR), _i(faceR), nleaf.value, key.ctypes.data_as(_vp),
                                            _i(cnt), _d(box))
        else:
            self.h = L.fb_tree_build(mesh.h, depth, height, order, order_mom, sep)
        if not self.h:
            raise FabipbError(L.fb_host_error().decode())
        self.depth, self.height = depth, height
        self.a = {k: _arr(L.fb_tree_array, self.h, k) for k in TREE_KEYS}
        self.n = len(self.a["area"])
        self.ncube = len(self.a["cubeLevel"])
        self.nleaf = len(self.a["leafStart"])
        self.max_order = int(max(self.a["ordMom"][height:].max(), self.a["ordM2L"][height:].max()))
        self.mesh = mesh

    def __del__(self):
        if getattr(self, "h", None):
            host_lib().fb_tree_free(self.h)
            self.h = None

    def oracle_layout(self, kappa, eps1, eps2, qorder):
        a = self.a
        return dict(kappa=kappa, eps=eps2 / eps1, eps1=eps1, qorder=qorder,
                    vtx=a["vtx"].reshape(-1, 9), nrm=a["nrm"].reshape(-1, 3), area=a["area"],
                    depth=self.depth, height=self.height, maxOrder=self.max_order,
                    ordMom=a["ordMom"], ordM2L=a["ordM2L"], levStart=a["levStart"],
                    levCount=a["levCount"], cubeLevel=a["cubeLevel"],
                    cubeX=a["cubeX"].reshape(-1, 3), kidStart=a["kidStart"], nKids=a["nKids"],
                    leafStart=a["leafStart"], leafCount=a["leafCount"], nearOff=a["nearOff"],
                    nearSrc=a["nearSrc"], m2lSrc=a["m2lSrc"], m2lDst=a["m2lDst"])

    def charge_tree(self):
        """Charge octree over the mesh's charges (src/chargeTree.c:228-307)."""
        L = host_lib()
        h = L.fb_ctree_build(self.mesh.h, self.depth)
        try:
            keys = ["levStart", "levCount", "level", "kidStart", "nKids", "x", "eRad", "chgStart",
                    "nChgs", "chgIdx"]
            ct = {k: _arr(L.fb_ctree_array, h, k) for k in keys}
        finally:
            L.fb_ctree_free(h)
        ct["depth"] = self.depth
        ct["x"] = ct["x"].reshape(-1, 3)
        ct["order"] = np.array([L.fb_charge_tree_order(self.h, lev) for lev in range(self.depth + 1)],
                               np.int32)
        return ct


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise FabipbError("no GPU visible to torch")
    return torch


class Operator:
    """The device-resident BIE operator (fbh_create ... fbh_destroy)."""

    def __init__(self, tree, kappa=0.1257, eps1=1.0, eps2=80.0, qorder=1, stream=None):
        torch = _torch()
        L = hip_lib()
        if not L.fbh_available():
            raise FabipbError(_hip_err())
        a = tree.a
        self.tree, self.kappa, self.eps1, self.eps2, self.qorder = tree, kappa, eps1, eps2, qorder
        self.n = tree.n
        if stream is None:
            stream = torch.cuda.current_stream().cuda_stream
        self._keep = a
        t0 = time.perf_counter()
        self.h = L.fbh_create(kappa, eps1, eps2, qorder, tree.n, _d(a["vtx"]), _d(a["nrm"]),
                              _d(a["area"]), tree.depth, tree.height, _i(a["ordMom"]),
                              _i(a["ordM2L"]), tree.ncube, _i(a["levStart"]), _i(a["levCount"]),
                              _i(a["cubeLevel"]), _d(a["cubeX"]), _i(a["kidStart"]), _i(a["nKids"]),
                              _i(a["parent"]), tree.nleaf, _i(a["leafStart"]), _i(a["leafCount"]),
                              _i(a["nearOff"]), _i(a["nearSrc"]), _i(a["m2lOff"]), _i(a["m2lSrc"]),
                              C.c_void_p(stream))
        if not self.h:
            raise FabipbError(_hip_err())
        self.setup_s = time.perf_counter() - t0

    def __del__(self):
        if getattr(self, "h", None):
            hip_lib().fbh_destroy(self.h)
            self.h = None

    def _ck(self, ok):
        if not ok:
            raise FabipbError(_hip_err())

    def stats(self):
        out = (C.c_longlong * 16)()
        self._ck(hip_lib().fbh_stats(self.h, out))
        keys = ["n", "near_interactions", "m2l_pairs", "nmom_max", "nmom_leaf", "device_bytes",
                "links", "cubes", "near_stored", "near_fly_tiles", "near_tiles", "special_pairs", "m2l_grid_steps"]
        return dict(zip(keys, list(out)))

    # host-array calls
    def matvec(self, x, alpha=1.0, beta=0.0, y=None):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.zeros(2 * self.n) if y is None else np.array(y, dtype=np.float64)
        self._ck(hip_lib().fbh_matvec(self.h, alpha, _d(x), beta, _d(y)))
        return y

    def near_apply(self, x, alpha=1.0, pot=None):
        x = np.ascontiguousarray(x, dtype=np.float64)
        pot = np.zeros(2 * self.n) if pot is None else np.array(pot, dtype=np.float64)
        self._ck(hip_lib().fbh_near_apply(self.h, alpha, _d(x), _d(pot)))
        return pot

    def far_apply(self, x, alpha=1.0, beta=0.0, pot=None):
        x = np.ascontiguousarray(x, dtype=np.float64)
        pot = np.zeros(2 * self.n) if pot is None else np.array(pot, dtype=np.float64)
        self._ck(hip_lib().fbh_far_apply(self.h, alpha, _d(x), beta, _d(pot)))
        return pot

    def expansions(self):
        st = self.stats()
        nm = st["nmom_max"]
        mom = np.zeros(self.tree.ncube * 2 * nm)
        lec = np.zeros(self.tree.ncube * 4 * nm)
        self._ck(hip_lib().fbh_expansions(self.h, _d(mom), _d(lec)))
        return mom, lec

    def q2m_matrices(self):
        nm = self.stats()["nmom_leaf"]
        q0 = np.zeros(self.n * nm)
        q1 = np.zeros(self.n * nm)
        self._ck(hip_lib().fbh_q2m_matrices(self.h, _d(q0), _d(q1)))
        return q0, q1

    def near_coeff(self, x, y):
        k = np.zeros(4)
        self._ck(hip_lib().fbh_near_coeff(self.h, int(x), int(y), _d(k)))
        return k

    def set_rhs_factor(self, fac):
        """Scale of every right-hand side (the fac of gpuSetupRHS); 0 restores 1/(4 pi eps1)."""
        self._ck(hip_lib().fbh_set_rhs_factor(self.h, float(fac)))

    def pair_coeffs(self, x, y):
        """KER[4] = panelIA0(X = panel x[p], Y = panel y[p]) of listed pairs: array (nPairs, 4)."""
        x = np.ascontiguousarray(x, dtype=np.int32)
        y = np.ascontiguousarray(y, dtype=np.int32)
        k = np.zeros((4, len(x)))
        self._ck(hip_lib().fbh_pair_coeffs(self.h, len(x), _i(x), _i(y), _d(k[0]), _d(k[1]), _d(k[2]), _d(k[3])))
        return k.T.copy()

    def rhs(self, pos, chr_):
        pos = np.ascontiguousarray(pos, dtype=np.float64)
        chr_ = np.ascontiguousarray(chr_, dtype=np.float64)
        b = np.zeros(2 * self.n)
        self._ck(hip_lib().fbh_rhs(self.h, len(chr_), _d(pos), _d(chr_), _d(b)))
        return b

    # device-tensor calls (torch float64 cuda tensors of length 2n)
    def matvec_dev(self, x, y, alpha=1.0, beta=0.0):
        self._ck(hip_lib().fbh_matvec_dev(self.h, alpha, C.c_void_p(x.data_ptr()), beta,
                                          C.c_void_p(y.data_ptr())))

    def set_direct(self, on):
        """-r=1/2: the dense direct operator (every pair's panelIA0) in place of the FMM."""
        self._ck(hip_lib().fbh_set_direct(self.h, int(on)))

    def set_precond(self, mode):
        """GMRES preconditioner (src/fabipb.c:1289-1309): 3 diagonal; 0, 1, 2 the
        leaf-parent block LU (src/precond_fmm.c), built as explicit block inverses."""
        self._ck(hip_lib().fbh_set_precond(self.h, mode))

    def psolve_dev(self, x, y):
        self._ck(hip_lib().fbh_psolve_dev(self.h, C.c_void_p(y.data_ptr()), C.c_void_p(x.data_ptr())))

    def enable_fp32(self):
        """FP32 copies of the operator for the precision form (fbh_enable_fp32)."""
        self._ck(hip_lib().fbh_enable_fp32(self.h))

    def matvec32_dev(self, x, y, alpha=1.0, beta=0.0):
        self._ck(hip_lib().fbh_matvec32_dev(self.h, alpha, C.c_void_p(x.data_ptr()), beta,
                                            C.c_void_p(y.data_ptr())))

    def matvec32(self, x, alpha=1.0, beta=0.0, y=None):
        """Host-array convenience: y = alpha (D - K32) x + beta y on the FP32 operator."""
        torch = _torch()
        xd = torch.tensor(np.ascontiguousarray(x, dtype=np.float64), device="cuda")
        yd = (torch.tensor(np.ascontiguousarray(y, dtype=np.float64), device="cuda") if y is not None
              else torch.zeros_like(xd))
        self.matvec32_dev(xd, yd, alpha, beta if y is not None else 0.0)
        return yd.cpu().numpy()

    def gmres_refined_dev(self, b, x, restart=30, tol=1e-6, maxit=100, max_outer=10):
        it, res, info, outer, mv64 = C.c_int(0), C.c_double(0.0), C.c_int(0), C.c_int(0), C.c_int(0)
        hist = np.zeros(max_outer)
        self._ck(hip_lib().fbh_gmres_refined_dev(
            self.h, C.c_void_p(b.data_ptr()), C.c_void_p(x.data_ptr()), restart, tol, maxit, max_outer,
            C.byref(it), C.byref(res), _d(hist), C.byref(info), C.byref(outer), C.byref(mv64)))
        return dict(iters=it.value, resid=res.value, info=info.value, outer=outer.value, mv64=mv64.value,
                    hist=hist[:outer.value])

    def gmres_refined_energy_dev(self, b, x, restart=30, tol=1e-6, maxit=100, max_outer=10, theta=0.8):
        """gmres_refined_dev, then the precision form's energy of the resident charges
        for the solution, the walk overlapped with the last FP64 residual product."""
        it, res, info, outer, mv64 = C.c_int(0), C.c_double(0.0), C.c_int(0), C.c_int(0), C.c_int(0)
        e = C.c_double(0.0)
        hist = np.zeros(max_outer)
        self._ck(hip_lib().fbh_gmres_refined_energy_dev(
            self.h, C.c_void_p(b.data_ptr()), C.c_void_p(x.data_ptr()), restart, tol, maxit, max_outer, theta,
            C.byref(it), C.byref(res), _d(hist), C.byref(info), C.byref(outer), C.byref(mv64), C.byref(e)))
        return dict(iters=it.value, resid=res.value, info=info.value, outer=outer.value, mv64=mv64.value,
                    hist=hist[:outer.value], energy_raw=e.value)

    def matvec_t_dev(self, x, y, alpha=1.0, beta=0.0):
        """y = alpha A^T x + beta y (the adjoint operator)."""
        self._ck(hip_lib().fbh_matvec_t_dev(self.h, alpha, C.c_void_p(x.data_ptr()), beta,
                                            C.c_void_p(y.data_ptr())))

    def gmres_t_dev(self, b, x, restart=30, tol=1e-6, maxit=100):
        it, res, info = C.c_int(0), C.c_double(0.0), C.c_int(0)
        hist = np.zeros(2 * maxit + 2)   # one entry per iteration and one per restart cycle
        self._ck(hip_lib().fbh_gmres_t_dev(self.h, C.c_void_p(b.data_ptr()), C.c_void_p(x.data_ptr()),
                                           restart, tol, maxit, C.byref(it), C.byref(res), _d(hist),
                                           C.byref(info)))
        return dict(iters=it.value, resid=res.value, info=info.value, hist=hist)

    def energy_direct_dev(self, pos, chr_, sgm, g=None):
        """Raw direct energy g . sgm (times 2 pi 332.0716 for kcal/mol); g (2n tensor) optional out."""
        pos = np.ascontiguousarray(pos, dtype=np.float64)
        chr_ = np.ascontiguousarray(chr_, dtype=np.float64)
        e = C.c_double(0.0)
        self._ck(hip_lib().fbh_energy_direct_dev(
            self.h, len(chr_), _d(pos), _d(chr_), C.c_void_p(sgm.data_ptr()),
            C.c_void_p(g.data_ptr()) if g is not None else None, C.byref(e)))
        return e.value

    def gradient_dev(self, pos, chr_, sgm, lam, restart=30, tol=1e-6, maxit=100):
        """Adjoint gradient of the raw direct energy: (dE/dq [nChar], dE/dkappa, solve info); lam in/out."""
        pos = np.ascontiguousarray(pos, dtype=np.float64)
        chr_ = np.ascontiguousarray(chr_, dtype=np.float64)
        dq = np.zeros(len(chr_))
        dk, res = C.c_double(0.0), C.c_double(0.0)
        it, info = C.c_int(0), C.c_int(0)
        self._ck(hip_lib().fbh_gradient_dev(
            self.h, len(chr_), _d(pos), _d(chr_), C.c_void_p(sgm.data_ptr()), C.c_void_p(lam.data_ptr()),
            restart, tol, maxit, _d(dq), C.byref(dk), C.byref(it), C.byref(res), C.byref(info)))
        return dq, dk.value, dict(iters=it.value, resid=res.value, info=info.value)

    def gradient_tree_dev(self, ct, pos, chr_, sgm, lam, theta=0.8, rhs_tree=False, restart=30, tol=1e-6,
                          maxit=100, lam_from_rhs=False):
        """Exact gradient of the raw panel-tree energy (fbh_gradient_tree_dev): (dE/dq [nChar],
        dE/dkappa, adjoint solve info); rhs_tree: the solve's right-hand side was the
        charge-tree one; lam: the adjoint solution out, its initial guess in (or the adjoint's
        right-hand side with lam_from_rhs)."""
        pos = np.ascontiguousarray(pos, dtype=np.float64)
        chr_ = np.ascontiguousarray(chr_, dtype=np.float64)
        k = {q: np.ascontiguousarray(v) for q, v in ct.items() if isinstance(v, np.ndarray)}
        dq = np.zeros(len(chr_))
        dk, res = C.c_double(0.0), C.c_double(0.0)
        it, info = C.c_int(0), C.c_int(0)
        self._ck(hip_lib().fbh_gradient_tree_dev(
            self.h, len(chr_), _d(pos), _d(chr_), len(k["level"]), _i(k["level"]), _d(k["x"]), _d(k["eRad"]),
            _i(k["kidStart"]), _i(k["nKids"]), _i(k["chgStart"]), _i(k["nChgs"]), _i(k["chgIdx"]), _i(k["order"]),
            theta, int(bool(rhs_tree)), C.c_void_p(sgm.data_ptr()), C.c_void_p(lam.data_ptr()), int(bool(lam_from_rhs)),
            restart, tol, maxit, _d(dq), C.byref(dk), C.byref(it), C.byref(res), C.byref(info)))
        return dq, dk.value, dict(iters=it.value, resid=res.value, info=info.value)

    def tree_adjoint_dev(self, ct, pos, chr_, w, kind=0, theta=0.8):
        """d/dq at fixed densities of the panel-tree energy of w (kind 0) or of
        w . (charge-tree right-hand side) (kind 1), by the transposed walk."""
        pos = np.ascontiguousarray(pos, dtype=np.float64)
        chr_ = np.ascontiguousarray(chr_, dtype=np.float64)
        k = {q: np.ascontiguousarray(v) for q, v in ct.items() if isinstance(v, np.ndarray)}
        dq = np.zeros(len(chr_))
        self._ck(hip_lib().fbh_tree_adjoint_dev(
            self.h, len(chr_), _d(pos), _d(chr_), len(k["level"]), _i(k["level"]), _d(k["x"]), _d(k["eRad"]),
            _i(k["kidStart"]), _i(k["nKids"]), _i(k["chgStart"]), _i(k["nChgs"]), _i(k["chgIdx"]), _i(k["order"]),
            theta, int(kind), C.c_void_p(w.data_ptr()), _d(dq)))
        return dq

    def rhs_dev(self, pos, chr_, b):
        pos = np.ascontiguousarray(pos, dtype=np.float64)
        chr_ = np.ascontiguousarray(chr_, dtype=np.float64)
        self._ck(hip_lib().fbh_rhs_dev(self.h, len(chr_), _d(pos), _d(chr_),
                                       C.c_void_p(b.data_ptr())))

    def rhs_dev_noalloc(self, pos, chr_, b):
        """rhs_dev with the charges staged in the library's static device array (no allocation)."""
        pos = np.ascontiguousarray(pos, dtype=np.float64)
        chr_ = np.ascontiguousarray(chr_, dtype=np.float64)
        self._ck(hip_lib().fbh_rhs_dev_noalloc(self.h, len(chr_), _d(pos), _d(chr_), C.c_void_p(b.data_ptr())))

    def gmres_dev(self, b, x, restart=30, tol=1e-6, maxit=100):
        it, res, info = C.c_int(0), C.c_double(0.0), C.c_int(0)
        hist = np.zeros(2 * maxit + 2)
        self._ck(hip_lib().fbh_gmres_dev(self.h, C.c_void_p(b.data_ptr()), C.c_void_p(x.data_ptr()),
                                         restart, tol, maxit, C.byref(it), C.byref(res), _d(hist),
                                         C.byref(info)))
        return dict(iters=it.value, resid=res.value, info=info.value, hist=hist)

    def energy_dev(self, ct, pos, chr_, sigma, theta=0.8):
        pos = np.ascontiguousarray(pos, dtype=np.float64)
        chr_ = np.ascontiguousarray(chr_, dtype=np.float64)
        keep = {k: np.ascontiguousarray(v) for k, v in ct.items() if isinstance(v, np.ndarray)}
        e = C.c_double(0.0)
        self._ck(hip_lib().fbh_energy_dev(
            self.h, len(chr_), _d(pos), _d(chr_), len(keep["level"]), _i(keep["level"]),
            _d(keep["x"]), _d(keep["eRad"]), _i(keep["kidStart"]), _i(keep["nKids"]),
            _i(keep["chgStart"]), _i(keep["nChgs"]), _i(keep["chgIdx"]), _i(keep["order"]),
            theta, C.c_void_p(sigma.data_ptr()), C.byref(e)))
        return e.value

    def rhs_tree_dev(self, ct, pos, chr_, b, theta=0.8):
        """Charge-tree right-hand side into the device tensor b (setupRHSTree)."""
        pos = np.ascontiguousarray(pos, dtype=np.float64)
        chr_ = np.ascontiguousarray(chr_, dtype=np.float64)
        keep = {k: np.ascontiguousarray(v) for k, v in ct.items() if isinstance(v, np.ndarray)}
        self._ck(hip_lib().fbh_rhs_tree_dev(
            self.h, len(chr_), _d(pos), _d(chr_), len(keep["level"]), _i(keep["level"]),
            _d(keep["x"]), _d(keep["eRad"]), _i(keep["kidStart"]), _i(keep["nKids"]),
            _i(keep["chgStart"]), _i(keep["nChgs"]), _i(keep["chgIdx"]), _i(keep["order"]),
            theta, C.c_void_p(b.data_ptr())))

    def set_charges(self, ct, pos, chr_):
        """Stages the charges and their charge tree on the device once (fbh_set_charges);
        rhs_res_dev, rhs_tree_res_dev and energy_res_dev then read them there."""
        pos = np.ascontiguousarray(pos, dtype=np.float64)
        chr_ = np.ascontiguousarray(chr_, dtype=np.float64)
        k = {q: np.ascontiguousarray(v) for q, v in ct.items() if isinstance(v, np.ndarray)}
        self._ck(hip_lib().fbh_set_charges(
            self.h, len(chr_), _d(pos), _d(chr_), len(k["level"]), _i(k["level"]), _d(k["x"]), _d(k["eRad"]),
            _i(k["kidStart"]), _i(k["nKids"]), _i(k["chgStart"]), _i(k["nChgs"]), _i(k["chgIdx"]), _i(k["order"])))

    def rhs_res_dev(self, b):
        """Direct right-hand side of the resident charges into the device tensor b."""
        self._ck(hip_lib().fbh_rhs_res_dev(self.h, C.c_void_p(b.data_ptr())))

    def rhs32_res_dev(self, b):
        """The precision form's direct right-hand side (FP32 kernel) of the resident charges."""
        self._ck(hip_lib().fbh_rhs32_res_dev(self.h, C.c_void_p(b.data_ptr())))

    def rhs_tree_res_dev(self, b, theta=0.8):
        """Charge-tree right-hand side of the resident charges into b."""
        self._ck(hip_lib().fbh_rhs_tree_res_dev(self.h, theta, C.c_void_p(b.data_ptr())))

    def energy_res_dev(self, sigma, theta=0.8, precision="fp64"):
        """Raw panel-tree energy of the resident charges for the density sigma
        (precision="fp32": the walk's far polynomials in FP32, the precision form's)."""
        e = C.c_double(0.0)
        f = hip_lib().fbh_energy32_res_dev if precision == "fp32" else hip_lib().fbh_energy_res_dev
        self._ck(f(self.h, theta, C.c_void_p(sigma.data_ptr()), C.byref(e)))
        return e.value

    def near_timing(self, on):
        """Starts (on=True, clearing) or stops event timing of every nearfield launch."""
        self._ck(hip_lib().fbh_near_timing(self.h, int(bool(on))))

    def near_timing_read(self):
        """{"fp64": (launches, total ms), "fp32": (...)} recorded since near_timing(True)."""
        cnt = (C.c_longlong * 2)()
        ms = np.zeros(2)
        self._ck(hip_lib().fbh_near_timing_read(self.h, cnt, _d(ms)))
        return {"fp64": (int(cnt[0]), float(ms[0])), "fp32": (int(cnt[1]), float(ms[1]))}

    def profile(self, x, y, iters=10, precision="fp64"):
        ms = np.zeros(9)
        f = hip_lib().fbh_profile32 if precision == "fp32" else hip_lib().fbh_profile
        self._ck(f(self.h, C.c_void_p(x.data_ptr()), C.c_void_p(y.data_ptr()), iters, _d(ms)))
        return dict(zip(["near", "q2m", "m2m", "m2l", "l2l", "l2p", "combine", "total", "far_chain"], ms))


class Problem:
    """Differentiation form: a fixed mesh and tree; E(q, kappa) through the direct
    right-hand side (or the charge-tree one, rhs="tree"), GMRES on the FP64 operator
    and the panel-tree energy the plain solve reports (energy="tree", theta 0.8:
    src/treecode.c:927-1017), or the direct energy (energy="direct": every charge
    taken directly), and the exact adjoint gradient of that energy
    (Operator.gradient_tree_dev or Operator.gradient_dev).  Operators are built per
    kappa and kept (the nearfield corrections and M2L tensors depend on kappa)."""

    def __init__(self, mesh, depth, order=6, sep=0.8, eps1=1.0, eps2=80.0, qorder=1, height=2, tol=1e-6,
                 restart=30, maxit=100, precond=3, energy="tree", rhs="direct", theta=0.8):
        if energy not in ("tree", "direct") or rhs not in ("direct", "tree"):
            raise ValueError(f"energy must be tree or direct and rhs direct or tree, not {energy!r}, {rhs!r}")
        if energy == "direct" and rhs == "tree":
            raise ValueError("the direct energy's gradient takes the direct right-hand side")
        self.mesh = mesh
        self.tree = Tree(mesh, depth, order=order, height=height, sep=sep)
        self.ct = self.tree.charge_tree()
        self.eps1, self.eps2, self.qorder = eps1, eps2, qorder
        self.tol, self.restart, self.maxit, self.precond = tol, restart, maxit, precond
        self.energy_kind, self.rhs_kind, self.theta = energy, rhs, theta
        self._ops = {}

    def operator(self, kappa):
        kappa = float(kappa)
        if kappa not in self._ops:
            op = Operator(self.tree, kappa, self.eps1, self.eps2, self.qorder)
            op.set_precond(self.precond)
            self._ops[kappa] = op
        return self._ops[kappa]

    def solve(self, q, kappa):
        """Solution u (device tensor) and the raw energy for charges q."""
        torch = _torch()
        op = self.operator(kappa)
        q = np.ascontiguousarray(q, dtype=np.float64)
        b = torch.empty(2 * self.tree.n, dtype=torch.float64, device="cuda")
        op.set_charges(self.ct, self.mesh.pos, q)
        if self.rhs_kind == "tree":
            op.rhs_tree_res_dev(b, theta=self.theta)
        else:
            op.rhs_res_dev(b)
        x = b.clone()
        g = op.gmres_dev(b, x, restart=self.restart, tol=self.tol, maxit=self.maxit)
        if g["info"] != 0:
            raise RuntimeError(f"forward GMRES did not converge: {g}")
        if self.energy_kind == "tree":
            e = op.energy_res_dev(x, theta=self.theta)
        else:
            e = op.energy_direct_dev(self.mesh.pos, q, x)
        return x, e, g

    def energy(self, q, kappa):
        """Solvation energy in kcal/mol."""
        return self.solve(q, kappa)[1] * ENERGY_SCALE

    def gradient(self, q, kappa, u=None, e=None):
        """(E, dE/dq, dE/dkappa) in kcal/mol units; u: a solution already at hand and e
        its raw energy (recomputed from u when not given)."""
        torch = _torch()
        q = np.ascontiguousarray(q, dtype=np.float64)
        op = self.operator(kappa)
        if u is None:
            u, e, _ = self.solve(q, kappa)
        elif e is None:
            op.set_charges(self.ct, self.mesh.pos, q)
            e = (op.energy_res_dev(u, theta=self.theta) if self.energy_kind == "tree"
                 else op.energy_direct_dev(self.mesh.pos, q, u))
        lam = torch.zeros_like(u)
        if self.energy_kind == "tree":
            dq, dk, info = op.gradient_tree_dev(self.ct, self.mesh.pos, q, u, lam, theta=self.theta,
                                                rhs_tree=self.rhs_kind == "tree", restart=self.restart,
                                                tol=self.tol, maxit=self.maxit)
        else:
            dq, dk, info = op.gradient_dev(self.mesh.pos, q, u, lam, restart=self.restart, tol=self.tol,
                                           maxit=self.maxit)
        if info["info"] != 0:
            raise RuntimeError(f"adjoint GMRES did not converge: {info}")
        return e * ENERGY_SCALE, dq * ENERGY_SCALE, dk * ENERGY_SCALE


def solvation_energy(problem, q, kappa):
    """E(q, kappa) as a torch function with the adjoint gradient as its backward:
    q a float64 tensor of charges, kappa a float64 scalar tensor."""
    torch = _torch()

    class _SolvationEnergy(torch.autograd.Function):
        @staticmethod
        def forward(ctx, qt, kt):
            u, e, _ = problem.solve(qt.detach().cpu().numpy(), float(kt))
            ctx.u, ctx.q, ctx.kappa = u, qt.detach().cpu().numpy(), float(kt)
            return torch.tensor(e * ENERGY_SCALE, dtype=torch.float64)

        @staticmethod
        def backward(ctx, gE):
            _, dq, dk = problem.gradient(ctx.q, ctx.kappa, u=ctx.u)
            return gE * torch.tensor(dq, dtype=torch.float64), gE * torch.tensor(dk, dtype=torch.float64)

    return _SolvationEnergy.apply(q, kappa)


def solve(base, depth, order=6, sep=0.8, kappa=0.1257, eps1=1.0, eps2=80.0, qorder=1, tol=1e-6,
          restart=30, maxit=100, height=2, x0="auto", precision="fp64", op=None, tree=None, mesh=None,
          precond=3, direct=False, rhs="auto"):
    """One solve (src/fabipb.c:1067-1210): returns energies, iterations and times.
    precision "fp64" is the plain form (GMRES on the FP64 operator); "fp32" is the
    precision form (the direct right-hand side's kernel in FP32, refinement around
    GMRES on the FP32 operator, FP64 residuals, the energy in FP64).
    precond is the -P mode (3 diagonal; 0, 1, 2 block LU); direct the -r=1/2 dense
    operator in place of the FMM; rhs "direct", "tree" or "auto" (the reference's
    pair-count policy, FABIPB_FORCE_TREE_RHS).  x0 "rhs" starts GMRES from b (the
    reference's default), "zero" from 0; "auto" is "rhs" for the plain form and
    "zero" for the precision form, whose first correction then solves A d = b
    without an initial residual matvec.  An existing mesh/tree/operator may be
    passed in to time the solve alone."""
    torch = _torch()
    t0 = time.perf_counter()
    mesh = mesh if mesh is not None else Mesh(base)
    tree = tree if tree is not None else Tree(mesh, depth, order=order, height=height, sep=sep)
    t1 = time.perf_counter()
    op = op if op is not None else Operator(tree, kappa, eps1, eps2, qorder)
    op.set_precond(precond)
    op.set_direct(direct)
    if precision == "fp32":
        op.enable_fp32()
    elif precision != "fp64":
        raise ValueError(f"precision must be fp64 or fp32, not {precision!r}")
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    n2 = 2 * tree.n
    b = torch.empty(n2, dtype=torch.float64, device="cuda")
    ct = tree.charge_tree()
    op.set_charges(ct, mesh.pos, mesh.chr)   # staged once for the right-hand side and the energy
    if rhs == "auto":   # src/fabipb_options.c:174-183
        rhs = "tree" if host_lib().fb_should_use_tree_rhs(tree.n, len(mesh.chr), 1) else "direct"
    if rhs == "tree":
        op.rhs_tree_res_dev(b)
    elif precision == "fp32":
        op.rhs32_res_dev(b)
    else:
        op.rhs_res_dev(b)
    if x0 == "auto":
        x0 = "zero" if precision == "fp32" else "rhs"
    x = b.clone() if x0 == "rhs" else torch.zeros_like(b)
    torch.cuda.synchronize()
    t3 = time.perf_counter()
    if precision == "fp32":
        g = op.gmres_refined_dev(b, x, restart=restart, tol=tol, maxit=maxit)
    else:
        g = op.gmres_dev(b, x, restart=restart, tol=tol, maxit=maxit)
    torch.cuda.synchronize()
    t4 = time.perf_counter()
    raw = op.energy_res_dev(x, precision=precision)
    torch.cuda.synchronize()
    t5 = time.perf_counter()
    return dict(energy=raw * ENERGY_SCALE, energy_raw=raw, iters=g["iters"], resid=g["resid"],
                info=g["info"], hist=g["hist"][:g["iters"] + 2], n=tree.n, precision=precision,
                outer=g.get("outer", 0), mv64=g.get("mv64", g["iters"] + 1),
                t_tree=t1 - t0, t_setup=t2 - t1, t_rhs=t3 - t2, t_gmres=t4 - t3, t_energy=t5 - t4,
                t_solve=t5 - t2, x=x.cpu().numpy(), op=op, tree=tree, mesh=mesh)
