"""Python access to libapcloud (the MI355X-native AP-Cloud solve-and-move path)
through its C ABI (include/apcloud.h).  The product path is the HIP library;
this module only marshals arrays.  Loading fails loudly when the library has
not been built.
"""
import ctypes as C
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libapcloud.so")

_D = C.POINTER(C.c_double)
_I = C.POINTER(C.c_int)
_U64 = C.POINTER(C.c_uint64)

APC_OK = 0
ERRORS = {
    -1: "APC_ERR_ARG", -2: "APC_ERR_HIP", -3: "APC_ERR_NOMEM", -4: "APC_ERR_NEIGH_MANY",
    -5: "APC_ERR_NEIGH_FEW", -6: "APC_ERR_NEIGH_INVALID", -7: "APC_ERR_RANK",
    -8: "APC_ERR_BALANCE", -9: "APC_ERR_UNSUPPORTED", -10: "APC_ERR_NO_DEVICE",
}

# C ABI entry points declared in include/apcloud.h
EXPORTS = [
    "apc_last_error", "apc_version", "apc_device_count", "apc_create", "apc_destroy",
    "apc_set_particles", "apc_get_particles", "apc_octree_build", "apc_get_sorted",
    "apc_octree_level", "apc_select", "apc_get_comp_sizes", "apc_get_comp", "apc_neighbors",
    "apc_solver_set_neighbors", "apc_solver_set_constants", "apc_build_system",
    "apc_solver_solve", "apc_solver_amg_levels", "apc_solver_amg_level", "apc_differentiate",
    "apc_get_fields", "apc_set_gradients", "apc_interp_kick", "apc_drift", "apc_max_density",
    "apc_step", "apc_kernel_times", "apc_reset_kernel_times", "apc_set_timing", "apc_sync",
    "apc_stream_copy", "apc_freeze_topology", "apc_set_state", "apc_checkpoint",
    "apc_adjoint_seed_cic", "apc_adjoint_cic_weights", "apc_adjoint_get", "apc_adjoint_set", "apc_adjoint_drift",
    "apc_adjoint_kick", "apc_adjoint_fields", "apc_gradient_adjoint", "apc_solve_adjoint",
    "apc_batch_create", "apc_batch_destroy", "apc_batch_size", "apc_batch_box",
    "apc_batch_last_error", "apc_batch_set_particles", "apc_batch_run",
    "apc_morton_sort_noalloc", "apc_set_comp", "apc_solver_set_fit_threshold",
    "apc_fit_conditions", "apc_create_packed", "apc_pack_set_masses", "apc_pack_rows",
    "apc_pack_solve_info", "apc_set_knob", "apc_get_knob",
]

# launch-shape knobs (APC_KNOB_* in include/apcloud.h), in enum order
KNOBS = ["search_chunk", "mdot_blocks", "maxpy_blocks", "seg_blocks", "amg_g8_rows",
         "amg_g4_rows", "amg_rg8_rows", "amg_nopre", "amg_wlevels"]


class Params(C.Structure):
    _fields_ = [
        ("bbox", C.c_double * 6), ("depth", C.c_int), ("bc", C.c_int), ("c", C.c_double),
        ("order", C.c_int), ("mindepth", C.c_int), ("maxdepth", C.c_int),
        ("leaf_capacity", C.c_int), ("lm", C.c_int), ("dirichlet_point", C.c_int),
        ("boundary_at_faces", C.c_int), ("bv_kind", C.c_int), ("bv_mass", C.c_double),
        ("bv_radius", C.c_double), ("bv_center", C.c_double * 3), ("bv_G", C.c_double),
    ]


class KspOpts(C.Structure):
    _fields_ = [("rtol", C.c_double), ("atol", C.c_double), ("dtol", C.c_double),
                ("max_it", C.c_int), ("restart", C.c_int), ("monitor", C.c_int),
                ("precision", C.c_int), ("pc_type", C.c_int)]


class SolveInfo(C.Structure):
    _fields_ = [("its_rho", C.c_int), ("reason_rho", C.c_int), ("rnorm_rho", C.c_double),
                ("its_phi", C.c_int), ("reason_phi", C.c_int), ("rnorm_phi", C.c_double),
                ("refine_rho", C.c_int), ("refine_phi", C.c_int), ("nc", C.c_int),
                ("nv", C.c_int), ("nb", C.c_int), ("fit_qr_rows", C.c_int),
                ("fit_max_cond", C.c_double)]


class CompView(C.Structure):
    _fields_ = [(n, _D) for n in ("x", "y", "z", "cell_length")] + \
               [(n, _I) for n in ("np", "first", "type", "first_neigh", "number_neigh", "neigh24")] + \
               [(n, _D) for n in ("density", "potential", "p_x", "p_y", "p_z", "p_xx", "p_yy",
                                  "p_zz", "p_xy", "p_xz", "p_yz")]


_lib = None


def lib():
    """The loaded libapcloud.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libapcloud.so is not built (run __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        L.apc_last_error.restype = C.c_char_p
        L.apc_last_error.argtypes = [C.c_void_p]
        L.apc_version.restype = C.c_char_p
        L.apc_device_count.argtypes = [_I]
        L.apc_create.argtypes = [C.POINTER(Params), C.POINTER(C.c_void_p)]
        L.apc_destroy.argtypes = [C.c_void_p]
        L.apc_create_packed.argtypes = [C.POINTER(Params), C.c_int, _I, C.POINTER(C.c_void_p)]
        L.apc_pack_set_masses.argtypes = [C.c_void_p, _D]
        L.apc_pack_rows.argtypes = [C.c_void_p, _I, _I]
        L.apc_pack_solve_info.argtypes = [C.c_void_p, _I, _I, _I, _I, _D, _D]
        L.apc_set_knob.argtypes = [C.c_void_p, C.c_int, C.c_int]
        L.apc_get_knob.argtypes = [C.c_void_p, C.c_int, _I]
        L.apc_set_particles.argtypes = [C.c_void_p, C.c_int] + [_D] * 6
        L.apc_get_particles.argtypes = [C.c_void_p] + [_D] * 9
        L.apc_octree_build.argtypes = [C.c_void_p]
        L.apc_get_sorted.argtypes = [C.c_void_p, _U64, _I]
        L.apc_octree_level.argtypes = [C.c_void_p, C.c_int, _I, _U64, _I, _I]
        L.apc_select.argtypes = [C.c_void_p, C.c_double, C.c_int, C.c_int, C.c_int, _I]
        L.apc_get_comp_sizes.argtypes = [C.c_void_p, _I, _I, _I, _I]
        L.apc_get_comp.argtypes = [C.c_void_p, C.POINTER(CompView)]
        L.apc_neighbors.argtypes = [C.c_void_p, _I, _D]
        L.apc_solver_set_fit_threshold.argtypes = [C.c_void_p, C.c_double]
        L.apc_fit_conditions.argtypes = [C.c_void_p, _D, _I, _D]
        L.apc_set_comp.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int] + [_D] * 4 + \
            [_I] * 5 + [C.c_int]
        L.apc_solver_set_neighbors.argtypes = [C.c_void_p, C.c_int]
        L.apc_solver_set_constants.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_double]
        L.apc_build_system.argtypes = [C.c_void_p, _I] + [_D] * 6
        L.apc_solver_solve.argtypes = [C.c_void_p, C.POINTER(KspOpts), C.POINTER(KspOpts),
                                       C.POINTER(SolveInfo)]
        L.apc_differentiate.argtypes = [C.c_void_p, _D]
        L.apc_get_fields.argtypes = [C.c_void_p, _D, _D, _D, _D]
        L.apc_set_gradients.argtypes = [C.c_void_p, _D, _D]
        L.apc_interp_kick.argtypes = [C.c_void_p, C.c_double, C.c_double, _D]
        L.apc_drift.argtypes = [C.c_void_p, C.c_double, C.c_double, _D]
        L.apc_max_density.argtypes = [C.c_void_p, _D]
        L.apc_step.argtypes = [C.c_void_p] + [C.c_double] * 6 + [C.POINTER(KspOpts)] * 2 + \
            [_D, _D, C.POINTER(SolveInfo), _D]
        L.apc_kernel_times.argtypes = [C.c_void_p, C.c_char_p, C.c_int, _D, _I, _D, C.c_int]
        L.apc_batch_create.argtypes = [C.POINTER(Params), C.c_int, C.c_int, C.POINTER(C.c_void_p)]
        L.apc_batch_destroy.argtypes = [C.c_void_p]
        L.apc_batch_size.argtypes = [C.c_void_p]
        L.apc_batch_box.argtypes = [C.c_void_p, C.c_int]
        L.apc_batch_box.restype = C.c_void_p
        L.apc_batch_last_error.argtypes = [C.c_void_p]
        L.apc_batch_last_error.restype = C.c_char_p
        L.apc_batch_set_particles.argtypes = [C.c_void_p, C.c_int, C.c_int] + [_D] * 6
        L.apc_morton_sort_noalloc.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, _D,
                                              C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
        L.apc_batch_run.argtypes = [C.c_void_p, C.c_int, _D, _D, C.POINTER(KspOpts),
                                    C.POINTER(KspOpts), C.POINTER(SolveInfo), _I]
        L.apc_stream_copy.argtypes = [C.c_void_p, C.c_double, C.c_int, _D]
        L.apc_reset_kernel_times.argtypes = [C.c_void_p]
        L.apc_set_timing.argtypes = [C.c_void_p, C.c_int]
        L.apc_sync.argtypes = [C.c_void_p]
        _lib = L
    return _lib


def _p(a, t=_D):
    return a.ctypes.data_as(t)


class ApcError(RuntimeError):
    pass


def make_params(bbox=(0, 1, 0, 1, 0, 1), depth=20, bc=1, c=1.0, lm=17, **kw):
    p = Params()
    p.bbox[:] = list(bbox)
    p.depth = depth
    p.bc = bc
    p.c = c
    p.order = kw.get("order", 2)
    p.mindepth = kw.get("mindepth", 0)
    p.maxdepth = kw.get("maxdepth", depth)
    p.leaf_capacity = kw.get("leaf_capacity", 0)
    p.lm = lm
    p.dirichlet_point = kw.get("dirichlet_point", 1)
    p.boundary_at_faces = kw.get("boundary_at_faces", 0)
    p.bv_kind = kw.get("bv_kind", 0)
    p.bv_mass = kw.get("bv_mass", 1.0)
    p.bv_radius = kw.get("bv_radius", 0.5)
    p.bv_center[:] = list(kw.get("bv_center", (0.5, 0.5, 0.5)))
    p.bv_G = kw.get("bv_G", 1.0)
    return p


PC_TYPES = {"jacobi": 0, "amg": 1, "gamg": 1}


def ksp(rtol=1e-5, atol=1e-50, dtol=1e5, max_it=10000, restart=30, monitor=0, precision=0,
        pc="jacobi"):
    """KSP options (PETSc defaults); pc: "jacobi" (the reference's) or "amg"."""
    o = KspOpts()
    o.rtol, o.atol, o.dtol = rtol, atol, dtol
    o.max_it, o.restart, o.monitor, o.precision = max_it, restart, monitor, precision
    o.pc_type = PC_TYPES[pc] if isinstance(pc, str) else int(pc)
    return o


def device_count():
    n = C.c_int()
    lib().apc_device_count(C.byref(n))
    return n.value


class Batch:
    """Many independent boxes stepped together (apc_batch_run): one context per
    box, driven by a pool of host threads; `boxes[i]` is box i's Context."""

    def __init__(self, params, nbox, nthreads=0):
        self.L = lib()
        self.p = params
        h = C.c_void_p()
        r = self.L.apc_batch_create(C.byref(params), nbox, nthreads, C.byref(h))
        if r != 0:
            raise ApcError("apc_batch_create failed: %s" % ERRORS.get(r, r))
        self.h = h
        self.boxes = [Context(params, _handle=C.c_void_p(self.L.apc_batch_box(h, i)))
                      for i in range(nbox)]

    def close(self):
        if getattr(self, "h", None):
            for b in self.boxes:
                b.close()
            self.L.apc_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def set_particles(self, i, x, y, z, vx, vy, vz):
        arrs = [np.ascontiguousarray(a, np.float64) for a in (x, y, z, vx, vy, vz)]
        r = self.L.apc_batch_set_particles(self.h, i, len(arrs[0]), *[_p(a) for a in arrs])
        if r != 0:
            raise ApcError("apc_batch_set_particles failed (%s): %s"
                           % (ERRORS.get(r, r), self.L.apc_batch_last_error(self.h).decode()))
        self.boxes[i].n = len(arrs[0])

    def run(self, coef, c_dens, rho_opts, phi_opts=None):
        """coef: (steps, 5) rows of (c_pot, c_field, Cv, Cx, dt); c_dens per
        box.  Returns (last SolveInfo per box, status per box)."""
        phi_opts = rho_opts if phi_opts is None else phi_opts
        coef = np.ascontiguousarray(coef, np.float64).reshape(-1, 5)
        c_dens = np.ascontiguousarray(c_dens, np.float64)
        nbox = len(self.boxes)
        assert c_dens.shape == (nbox,)
        info = (SolveInfo * nbox)()
        status = np.zeros(nbox, np.int32)
        r = self.L.apc_batch_run(self.h, len(coef), _p(coef), _p(c_dens), C.byref(rho_opts),
                                 C.byref(phi_opts), info, _p(status, _I))
        if r != 0:
            raise ApcError("apc_batch_run failed (%s): %s"
                           % (ERRORS.get(r, r), self.L.apc_batch_last_error(self.h).decode()))
        return list(info), status


class Context:
    """One simulation box on the current GPU."""

    def __init__(self, params, _handle=None):
        self.L = lib()
        self.p = params
        self.owned = _handle is None
        if _handle is None:
            h = C.c_void_p()
            r = self.L.apc_create(C.byref(params), C.byref(h))
            if r != 0:
                raise ApcError("apc_create failed: %s" % ERRORS.get(r, r))
            _handle = h
        self.h = _handle
        self.n = 0

    def close(self):
        if getattr(self, "h", None):
            if self.owned:
                self.L.apc_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def _chk(self, r, what):
        if r != 0:
            raise ApcError("%s failed (%s): %s" % (what, ERRORS.get(r, r),
                                                   self.L.apc_last_error(self.h).decode()))

    def set_particles(self, x, y, z, vx=None, vy=None, vz=None):
        self.n = len(x)
        arrs = [np.ascontiguousarray(a, np.float64) if a is not None else None
                for a in (x, y, z, vx, vy, vz)]
        ptrs = [_p(a) if a is not None else None for a in arrs]
        self._chk(self.L.apc_set_particles(self.h, self.n, *ptrs), "apc_set_particles")

    def particles(self):
        out = [np.zeros(self.n) for _ in range(9)]
        self._chk(self.L.apc_get_particles(self.h, *[_p(a) for a in out]), "apc_get_particles")
        return out

    def octree_build(self):
        self._chk(self.L.apc_octree_build(self.h), "apc_octree_build")

    def sorted(self):
        k = np.zeros(self.n, np.uint64)
        i = np.zeros(self.n, np.int32)
        self._chk(self.L.apc_get_sorted(self.h, _p(k, _U64), _p(i, _I)), "apc_get_sorted")
        return k, i

    def octree_level(self, level):
        cnt = C.c_int()
        self._chk(self.L.apc_octree_level(self.h, level, C.byref(cnt), None, None, None),
                  "apc_octree_level")
        m = cnt.value
        key = np.zeros(m, np.uint64)
        first = np.zeros(m, np.int32)
        npp = np.zeros(m, np.int32)
        self._chk(self.L.apc_octree_level(self.h, level, C.byref(cnt), _p(key, _U64),
                                          _p(first, _I), _p(npp, _I)), "apc_octree_level")
        return key, first, npp

    def select(self, c=None, order=2, mindepth=0, maxdepth=None):
        c = self.p.c if c is None else c
        maxdepth = self.p.depth if maxdepth is None else maxdepth
        S = C.c_int()
        self._chk(self.L.apc_select(self.h, c, order, mindepth, maxdepth, C.byref(S)), "apc_select")
        return S.value

    def comp(self, fields=False):
        nc, nv, nb, nn = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        self._chk(self.L.apc_get_comp_sizes(self.h, C.byref(nc), C.byref(nv), C.byref(nb),
                                            C.byref(nn)), "apc_get_comp_sizes")
        S = nc.value + nv.value + nb.value
        M = nc.value + nv.value
        d = {k: np.zeros(S) for k in ("x", "y", "z", "cell_length")}
        i = {k: np.zeros(S, np.int32) for k in ("np", "first", "type", "first_neigh", "number_neigh")}
        n24 = np.zeros(max(1, nn.value), np.int32)
        v = CompView()
        for k, a in d.items():
            setattr(v, k, _p(a))
        for k, a in i.items():
            setattr(v, k, _p(a, _I))
        v.neigh24 = _p(n24, _I)
        f = {}
        if fields:
            f = {k: np.zeros(M) for k in ("density", "potential", "p_x", "p_y", "p_z", "p_xx",
                                           "p_yy", "p_zz", "p_xy", "p_xz", "p_yz")}
            for k, a in f.items():
                setattr(v, k, _p(a))
        self._chk(self.L.apc_get_comp(self.h, C.byref(v)), "apc_get_comp")
        out = dict(nc=nc.value, nv=nv.value, nb=nb.value, x=d["x"], y=d["y"], z=d["z"],
                   h=d["cell_length"], np=i["np"], first=i["first"], type=i["type"],
                   fneigh=i["first_neigh"], nneigh=i["number_neigh"], neigh24=n24[:nn.value])
        out.update(f)
        return out

    def set_comp(self, nc, nv, nb, x, y, z, h, np_, neigh):
        """Caller-filled computational particles (apc_set_comp); neigh: one
        list of Neigh24 entries per row (nc + nv rows; boundary rows none)."""
        S = nc + nv + nb
        d = [np.ascontiguousarray(a, np.float64) for a in (x, y, z, h)]
        npp = np.ascontiguousarray(np_, np.int32)
        typ = np.array([0] * nc + [-1] * nv + [1] * nb, np.int32)
        cnt = np.array([len(neigh[i]) if i < nc + nv else 0 for i in range(S)], np.int32)
        first = np.concatenate([[0], np.cumsum(cnt)[:-1]]).astype(np.int32)
        flat = np.array([m for i in range(nc + nv) for m in neigh[i]] or [0], np.int32)
        self._chk(self.L.apc_set_comp(self.h, nc, nv, nb, *[_p(a) for a in d], _p(npp, _I),
                                      _p(typ, _I), _p(first, _I), _p(cnt, _I), _p(flat, _I),
                                      int(cnt.sum())), "apc_set_comp")

    def set_fit_threshold(self, cond_qr):
        self._chk(self.L.apc_solver_set_fit_threshold(self.h, float(cond_qr)),
                  "apc_solver_set_fit_threshold")

    def fit_conditions(self):
        """(per-row condition estimates, rows fitted by QR, largest estimate
        of the others) of the last system build."""
        cond = np.zeros(self.comp_rows())
        q, mx = C.c_int(), C.c_double()
        self._chk(self.L.apc_fit_conditions(self.h, _p(cond), C.byref(q), C.byref(mx)),
                  "apc_fit_conditions")
        return cond, q.value, mx.value

    def neighbors(self):
        nc, nv, nb, nn = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        self._chk(self.L.apc_get_comp_sizes(self.h, C.byref(nc), C.byref(nv), C.byref(nb), None),
                  "apc_get_comp_sizes")
        M = nc.value + nv.value
        lm = self.p.lm
        nbr = np.zeros(M * lm, np.int32)
        dist = np.zeros(M * lm)
        self._chk(self.L.apc_neighbors(self.h, _p(nbr, _I), _p(dist)), "apc_neighbors")
        return nbr.reshape(M, lm), dist.reshape(M, lm)

    def set_constants(self, c_dens, c_pot, c_field):
        self._chk(self.L.apc_solver_set_constants(self.h, c_dens, c_pot, c_field),
                  "apc_solver_set_constants")

    def build_system(self):
        nc, nv, nb = C.c_int(), C.c_int(), C.c_int()
        self._chk(self.L.apc_get_comp_sizes(self.h, C.byref(nc), C.byref(nv), C.byref(nb), None),
                  "apc_get_comp_sizes")
        M = nc.value + nv.value
        lm = self.p.lm
        cols = np.zeros(M * lm, np.int32)
        a1 = np.zeros(M * lm)
        a2 = np.zeros(M * lm)
        vec = [np.zeros(M) for _ in range(4)]
        self._chk(self.L.apc_build_system(self.h, _p(cols, _I), _p(a1), _p(a2),
                                          *[_p(v) for v in vec]), "apc_build_system")
        return dict(cols=cols.reshape(M, lm), a1=a1.reshape(M, lm), a2=a2.reshape(M, lm),
                    d1=vec[0], d2=vec[1], rhs1=vec[2], pbnd=vec[3])

    def solve(self, rho_opts, phi_opts=None):
        info = SolveInfo()
        phi_opts = rho_opts if phi_opts is None else phi_opts
        self._chk(self.L.apc_solver_solve(self.h, C.byref(rho_opts), C.byref(phi_opts),
                                          C.byref(info)), "apc_solver_solve")
        return info

    def fields(self):
        nc, nv, nb = C.c_int(), C.c_int(), C.c_int()
        self._chk(self.L.apc_get_comp_sizes(self.h, C.byref(nc), C.byref(nv), C.byref(nb), None),
                  "apc_get_comp_sizes")
        M = nc.value + nv.value
        rho, phi, g, H = np.zeros(M), np.zeros(M), np.zeros(3 * M), np.zeros(6 * M)
        self._chk(self.L.apc_get_fields(self.h, _p(rho), _p(phi), _p(g), _p(H)), "apc_get_fields")
        return dict(rho=rho, phi=phi, g=g.reshape(M, 3), H=H.reshape(M, 6))

    def differentiate(self, phi):
        phi = np.ascontiguousarray(phi, np.float64)
        self._chk(self.L.apc_differentiate(self.h, _p(phi)), "apc_differentiate")

    def set_gradients(self, g, H):
        g = np.ascontiguousarray(g, np.float64).ravel()
        H = np.ascontiguousarray(H, np.float64).ravel()
        self._chk(self.L.apc_set_gradients(self.h, _p(g), _p(H)), "apc_set_gradients")

    def interp_kick(self, cv, dt):
        out = C.c_double()
        self._chk(self.L.apc_interp_kick(self.h, cv, dt, C.byref(out)), "apc_interp_kick")
        return out.value

    def drift(self, cx, dt):
        out = C.c_double()
        self._chk(self.L.apc_drift(self.h, cx, dt, C.byref(out)), "apc_drift")
        return out.value

    def max_density(self):
        out = C.c_double()
        self._chk(self.L.apc_max_density(self.h, C.byref(out)), "apc_max_density")
        return out.value

    def step(self, c_dens, c_pot, c_field, cv, cx, dt, rho_opts, phi_opts=None):
        phi_opts = rho_opts if phi_opts is None else phi_opts
        mdv, mdx = C.c_double(), C.c_double()
        info = SolveInfo()
        timing = np.zeros(3)
        self._chk(self.L.apc_step(self.h, c_dens, c_pot, c_field, cv, cx, dt, C.byref(rho_opts),
                                  C.byref(phi_opts), C.byref(mdv), C.byref(mdx), C.byref(info),
                                  _p(timing)), "apc_step")
        return mdv.value, mdx.value, info, timing

    def amg_levels(self):
        """[(rows, nnz)] of the multigrid built by the last pc="amg" solve."""
        rows = np.zeros(32, np.int32)
        nnz = np.zeros(32, np.int64)
        n = C.c_int()
        self._chk(self.L.apc_solver_amg_levels(self.h, _p(rows, _I), nnz.ctypes.data_as(C.c_void_p),
                                               32, C.byref(n)), "apc_solver_amg_levels")
        return [(int(rows[k]), int(nnz[k])) for k in range(min(n.value, 32))]

    def amg_level(self, level):
        """(CSR scipy matrix of level >= 1 or None for level 0, aggregate map
        to the next level or None for the coarsest)."""
        import scipy.sparse as sp
        levels = self.amg_levels()
        n, nnz = levels[level]
        agg = np.zeros(n, np.int32) if level + 1 < len(levels) else None
        if level == 0:
            self._chk(self.L.apc_solver_amg_level(self.h, 0, None, None, None,
                                                  _p(agg, _I) if agg is not None else None),
                      "apc_solver_amg_level")
            return None, agg
        rp, col, val = np.zeros(n + 1, np.int32), np.zeros(nnz, np.int32), np.zeros(nnz)
        self._chk(self.L.apc_solver_amg_level(self.h, level, _p(rp, _I), _p(col, _I), _p(val),
                                              _p(agg, _I) if agg is not None else None),
                  "apc_solver_amg_level")
        return sp.csr_matrix((val, col, rp), shape=(n, n)), agg

    # ---- differentiation form (topology held fixed) ----
    def freeze_topology(self, on=True):
        self._chk(self.L.apc_freeze_topology(self.h, 1 if on else 0), "apc_freeze_topology")

    def set_state(self, x, y, z, vx, vy, vz):
        arrs = [np.ascontiguousarray(a, np.float64) for a in (x, y, z, vx, vy, vz)]
        self._chk(self.L.apc_set_state(self.h, *[_p(a) for a in arrs]), "apc_set_state")

    def checkpoint(self, slot, restore=False):
        self._chk(self.L.apc_checkpoint(self.h, slot, 1 if restore else 0), "apc_checkpoint")

    def adjoint_seed_cic(self, n):
        out = C.c_double()
        self._chk(self.L.apc_adjoint_seed_cic(self.h, n, C.byref(out)), "apc_adjoint_seed_cic")
        return out.value

    def adjoint_cic_weights(self, w=None):
        """Per-particle weights of the CIC loss (caller order); None: equal."""
        if w is None:
            self._chk(self.L.apc_adjoint_cic_weights(self.h, None), "apc_adjoint_cic_weights")
            return
        w = np.ascontiguousarray(w, np.float64)
        assert w.shape == (self.n,)
        self._chk(self.L.apc_adjoint_cic_weights(self.h, _p(w)), "apc_adjoint_cic_weights")

    def adjoint_get(self):
        out = [np.zeros(self.n) for _ in range(6)]
        self._chk(self.L.apc_adjoint_get(self.h, *[_p(a) for a in out]), "apc_adjoint_get")
        return out

    def adjoint_set(self, lx, ly, lz, lvx, lvy, lvz):
        arrs = [np.ascontiguousarray(a, np.float64) for a in (lx, ly, lz, lvx, lvy, lvz)]
        self._chk(self.L.apc_adjoint_set(self.h, *[_p(a) for a in arrs]), "apc_adjoint_set")

    def adjoint_drift(self, cx, dt):
        self._chk(self.L.apc_adjoint_drift(self.h, C.c_double(cx), C.c_double(dt)),
                  "apc_adjoint_drift")

    def adjoint_kick(self, cv, dt, accumulate=False, clear=False):
        self._chk(self.L.apc_adjoint_kick(self.h, C.c_double(cv), C.c_double(dt),
                                          1 if accumulate else 0, 1 if clear else 0),
                  "apc_adjoint_kick")

    def adjoint_fields(self):
        M = self.comp_rows()
        dg, dH = np.zeros(3 * M), np.zeros(6 * M)
        self._chk(self.L.apc_adjoint_fields(self.h, _p(dg), _p(dH)), "apc_adjoint_fields")
        return dg.reshape(M, 3), dH.reshape(M, 6)

    def gradient_adjoint(self, dg, dH):
        M = self.comp_rows()
        dg = np.ascontiguousarray(dg, np.float64).reshape(M * 3)
        dH = np.ascontiguousarray(dH, np.float64).reshape(M * 6)
        out = np.zeros(M)
        self._chk(self.L.apc_gradient_adjoint(self.h, _p(dg), _p(dH), _p(out)),
                  "apc_gradient_adjoint")
        return out

    def solve_adjoint(self, which, opts, rhs):
        rhs = np.ascontiguousarray(rhs, np.float64)
        out = np.zeros_like(rhs)
        its, reason = C.c_int(), C.c_int()
        self._chk(self.L.apc_solve_adjoint(self.h, which, C.byref(opts), _p(rhs), _p(out),
                                           C.byref(its), C.byref(reason)), "apc_solve_adjoint")
        return out, its.value, reason.value

    def comp_rows(self):
        nc, nv = C.c_int(), C.c_int()
        self._chk(self.L.apc_get_comp_sizes(self.h, C.byref(nc), C.byref(nv), None, None),
                  "apc_get_comp_sizes")
        return nc.value + nv.value

    def set_knob(self, name, value):
        self._chk(self.L.apc_set_knob(self.h, KNOBS.index(name), int(value)), "apc_set_knob")

    def knobs(self):
        out = {}
        for k, name in enumerate(KNOBS):
            v = C.c_int()
            self._chk(self.L.apc_get_knob(self.h, k, C.byref(v)), "apc_get_knob")
            out[name] = v.value
        return out

    def set_timing(self, on=True):
        self._chk(self.L.apc_set_timing(self.h, 1 if on else 0), "apc_set_timing")

    def reset_kernel_times(self):
        self._chk(self.L.apc_reset_kernel_times(self.h), "apc_reset_kernel_times")

    def kernel_times(self):
        buf = C.create_string_buffer(8192)
        avg = np.zeros(128)
        cnt = np.zeros(128, np.int32)
        byt = np.zeros(128)
        n = self.L.apc_kernel_times(self.h, buf, 8192, _p(avg), _p(cnt, _I), _p(byt), 128)
        names = buf.raw.split(b"\0")[:n]
        return {nm.decode(): (avg[q], int(cnt[q]), byt[q]) for q, nm in enumerate(names)}

    def stream_copy(self, nbytes, reps=10):
        out = C.c_double()
        self._chk(self.L.apc_stream_copy(self.h, float(nbytes), reps, C.byref(out)), "apc_stream_copy")
        return out.value

    def sync(self):
        self._chk(self.L.apc_sync(self.h), "apc_sync")


class PackedContext(Context):
    """Independent periodic boxes packed into one context
    (apc_create_packed): counts[b] particles of box b, given to
    set_particles box after box; every kernel of a step runs over all boxes
    in one launch."""

    def __init__(self, params, counts):
        counts = np.ascontiguousarray(counts, np.int32)
        L = lib()
        h = C.c_void_p()
        r = L.apc_create_packed(C.byref(params), len(counts), _p(counts, _I), C.byref(h))
        if r != 0:
            raise ApcError("apc_create_packed failed: %s" % ERRORS.get(r, r))
        super().__init__(params, _handle=h)
        self.owned = True
        self.counts = counts
        self.pstart = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)

    def set_masses(self, mass):
        m = np.ascontiguousarray(mass, np.float64)
        assert m.shape == (len(self.counts),)
        self._chk(self.L.apc_pack_set_masses(self.h, _p(m)), "apc_pack_set_masses")

    def rows(self):
        """(row_start[nbox + 1], comp_rows[nbox]) of the current selection."""
        nbox = len(self.counts)
        rs = np.zeros(nbox + 1, np.int32)
        nc = np.zeros(nbox, np.int32)
        self._chk(self.L.apc_pack_rows(self.h, _p(rs, _I), _p(nc, _I)), "apc_pack_rows")
        return rs, nc

    def solve_info(self):
        """Per box: (its_rho, its_phi, reason_rho, reason_phi, res_rho,
        res_phi) of the last solve."""
        nbox = len(self.counts)
        ints = [np.zeros(nbox, np.int32) for _ in range(4)]
        res = [np.zeros(nbox) for _ in range(2)]
        self._chk(self.L.apc_pack_solve_info(self.h, *[_p(a, _I) for a in ints],
                                             *[_p(a) for a in res]), "apc_pack_solve_info")
        return tuple(ints) + tuple(res)
