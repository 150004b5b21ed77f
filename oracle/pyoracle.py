"""TEST INFRASTRUCTURE ONLY: ctypes access to the oracle (liboracle.so) and,
where it was built, to oracle/_ref/libref.so (the reference's own octree.cpp /
mover.cpp compiled from its sources).  Only tests/, __graft_entry__.smoke()
and bench.py's cpu_baseline leg import this module.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_D = C.POINTER(C.c_double)
_I = C.POINTER(C.c_int)
_U64 = C.POINTER(C.c_uint64)


class OrcParams(C.Structure):
    _fields_ = [
        ("bbox", C.c_double * 6), ("depth", C.c_int), ("bc", C.c_int), ("c", C.c_double),
        ("order", C.c_int), ("mindepth", C.c_int), ("maxdepth", C.c_int), ("lm", C.c_int),
        ("dirichlet_point", C.c_int), ("boundary_at_faces", C.c_int), ("bv_kind", C.c_int),
        ("bv_mass", C.c_double), ("bv_radius", C.c_double), ("bv_center", C.c_double * 3),
        ("bv_G", C.c_double), ("units", C.c_int), ("omega", C.c_double),
        ("omega_lambda", C.c_double), ("omega_curv", C.c_double), ("G", C.c_double),
        ("unit_mass", C.c_double), ("unit_length", C.c_double),
        ("mass_per_particle", C.c_double), ("adaptive_dt", C.c_int), ("rtol", C.c_double),
        ("atol", C.c_double), ("dtol", C.c_double), ("max_it", C.c_int), ("restart", C.c_int),
        ("nthreads", C.c_int), ("leaf_capacity", C.c_int),
    ]


class OrcSolveInfo(C.Structure):
    _fields_ = [("its_rho", C.c_int), ("reason_rho", C.c_int), ("rnorm_rho", C.c_double),
                ("its_phi", C.c_int), ("reason_phi", C.c_int), ("rnorm_phi", C.c_double)]


def _p(a, t=_D):
    return a.ctypes.data_as(t)


_lib = None


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle not built: run make -C oracle")
        _lib = C.CDLL(path)
        L = _lib
        L.orc_new.restype = C.c_void_p
        L.orc_new.argtypes = [C.POINTER(OrcParams)]
        L.orc_free.argtypes = [C.c_void_p]
        for name in ["orc_set_particles"]:
            getattr(L, name).argtypes = [C.c_void_p, C.c_int] + [_D] * 6
        L.orc_get_particles.argtypes = [C.c_void_p] + [_D] * 9
        L.orc_key64.restype = C.c_uint64
        L.orc_key64.argtypes = [C.c_double] * 3 + [_D, C.c_int]
        L.orc_build_tree.argtypes = [C.c_void_p]
        L.orc_tree_info.argtypes = [C.c_void_p, _I, _I]
        L.orc_tree_sorted.argtypes = [C.c_void_p, _U64, _I]
        L.orc_tree_nodes.argtypes = [C.c_void_p, _U64] + [_I] * 5 + [_D, _D]
        L.orc_select.argtypes = [C.c_void_p, C.c_double, C.c_int, C.c_int, C.c_int]
        L.orc_comp_sizes.argtypes = [C.c_void_p, _I, _I, _I, _I]
        L.orc_comp_get.argtypes = [C.c_void_p] + [_D] * 4 + [_I] * 6
        L.orc_neighbors.argtypes = [C.c_void_p, _I, _D]
        L.orc_set_comp.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int] + [_D] * 4 + \
            [_I] * 4 + [C.c_int]
        L.orc_build_system.argtypes = [C.c_void_p, _I] + [_D] * 6
        L.orc_set_constants.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_double]
        L.orc_solve.argtypes = [C.c_void_p, C.POINTER(OrcSolveInfo)]
        L.orc_comp_fields.argtypes = [C.c_void_p] + [_D] * 4
        L.orc_set_potential.argtypes = [C.c_void_p, _D]
        L.orc_set_gradients.argtypes = [C.c_void_p, _D, _D]
        L.orc_differentiate.argtypes = [C.c_void_p]
        L.orc_move_velocity.restype = C.c_double
        L.orc_move_velocity.argtypes = [C.c_void_p, C.c_double, C.c_double]
        L.orc_move_position.restype = C.c_double
        L.orc_move_position.argtypes = [C.c_void_p, C.c_double, C.c_double]
        L.orc_run.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_double, C.c_int, _D, _I]
        L.orc_wlsq.argtypes = [C.c_int, C.c_int, C.c_int, _D, _D]
        L.orc_quicksort.argtypes = [C.c_int, _D, _I]
        L.orc_gmres_csr.argtypes = [C.c_int, _I, _I, _D, _D, _D, C.c_double, C.c_double,
                                    C.c_double, C.c_int, C.c_int, _I, _I, _D]
        L.orc_coefficients.argtypes = [C.POINTER(OrcParams), C.c_double] + [_D] * 5
    return _lib


def make_params(bbox=(0, 1, 0, 1, 0, 1), depth=20, bc=1, c=1.0, lm=17, **kw):
    p = OrcParams()
    p.bbox[:] = list(bbox)
    p.depth = depth
    p.bc = bc
    p.c = c
    p.order = kw.get("order", 2)
    p.mindepth = kw.get("mindepth", 0)
    p.maxdepth = kw.get("maxdepth", depth)
    p.lm = lm
    p.dirichlet_point = kw.get("dirichlet_point", 1)
    p.boundary_at_faces = kw.get("boundary_at_faces", 0)
    p.bv_kind = kw.get("bv_kind", 0)
    p.bv_mass = kw.get("bv_mass", 1.0)
    p.bv_radius = kw.get("bv_radius", 0.5)
    p.bv_center[:] = list(kw.get("bv_center", (0.5, 0.5, 0.5)))
    p.bv_G = kw.get("bv_G", 1.0)
    p.units = kw.get("units", 0)
    p.omega = kw.get("omega", 1.0)
    p.omega_lambda = kw.get("omega_lambda", 0.0)
    p.omega_curv = kw.get("omega_curv", 0.0)
    p.G = kw.get("G", 1.0)
    p.unit_mass = kw.get("unit_mass", 1.0)
    p.unit_length = kw.get("unit_length", 1.0)
    p.mass_per_particle = kw.get("mass_per_particle", 1.0)
    p.adaptive_dt = kw.get("adaptive_dt", 0)
    p.rtol = kw.get("rtol", 1e-5)
    p.atol = kw.get("atol", 1e-50)
    p.dtol = kw.get("dtol", 1e5)
    p.max_it = kw.get("max_it", 10000)
    p.restart = kw.get("restart", 30)
    p.nthreads = kw.get("nthreads", 0)
    p.leaf_capacity = kw.get("leaf_capacity", 0)
    return p


class Oracle:
    """One oracle simulation state (particles, tree, computational particles)."""

    def __init__(self, params):
        self.p = params
        self.L = lib()
        self.h = self.L.orc_new(C.byref(params))

    def __del__(self):
        if getattr(self, "h", None):
            self.L.orc_free(self.h)
            self.h = None

    def set_particles(self, x, y, z, vx=None, vy=None, vz=None):
        n = len(x)
        z0 = np.zeros(n)
        arrs = [np.ascontiguousarray(a if a is not None else z0, dtype=np.float64)
                for a in (x, y, z, vx, vy, vz)]
        self.n = n
        r = self.L.orc_set_particles(self.h, n, *[_p(a) for a in arrs])
        assert r == 0

    def particles(self):
        out = [np.zeros(self.n) for _ in range(9)]
        self.L.orc_get_particles(self.h, *[_p(a) for a in out])
        return out

    def build_tree(self):
        assert self.L.orc_build_tree(self.h) == 0

    def tree_sorted(self):
        k = np.zeros(self.n, np.uint64)
        i = np.zeros(self.n, np.int32)
        self.L.orc_tree_sorted(self.h, _p(k, _U64), _p(i, _I))
        return k, i

    def tree_nodes(self):
        ln = C.c_int()
        nn = np.zeros(self.p.depth + 1, np.int32)
        self.L.orc_tree_info(self.h, C.byref(ln), _p(nn, _I))
        n = ln.value
        key = np.zeros(n, np.uint64)
        ints = [np.zeros(n, np.int32) for _ in range(5)]
        lo = np.zeros(3 * n)
        hi = np.zeros(3 * n)
        self.L.orc_tree_nodes(self.h, _p(key, _U64), *[_p(a, _I) for a in ints], _p(lo), _p(hi))
        return dict(numnodes=nn, key=key, depth=ints[0], first=ints[1], count=ints[2],
                    fchild=ints[3], nchild=ints[4], lo=lo.reshape(-1, 3), hi=hi.reshape(-1, 3))

    def select(self, c=None, order=2, mindepth=0, maxdepth=None):
        c = self.p.c if c is None else c
        maxdepth = self.p.depth if maxdepth is None else maxdepth
        return self.L.orc_select(self.h, c, order, mindepth, maxdepth)

    def comp(self):
        nc, nv, nb, nn = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        assert self.L.orc_comp_sizes(self.h, C.byref(nc), C.byref(nv), C.byref(nb), C.byref(nn)) == 0
        S = nc.value + nv.value + nb.value
        d = [np.zeros(S) for _ in range(4)]
        i = [np.zeros(S, np.int32) for _ in range(5)]
        n24 = np.zeros(max(nn.value, 1), np.int32)
        self.L.orc_comp_get(self.h, *[_p(a) for a in d], *[_p(a, _I) for a in i], _p(n24, _I))
        return dict(nc=nc.value, nv=nv.value, nb=nb.value, x=d[0], y=d[1], z=d[2], h=d[3],
                    np=i[0], first=i[1], type=i[2], fneigh=i[3], nneigh=i[4],
                    neigh24=n24[:nn.value])

    def set_comp(self, nc, nv, nb, x, y, z, h, np_, neigh):
        """A caller-filled computational-particle set; neigh: one Neigh24 list
        per row (nc + nv rows)."""
        S = nc + nv + nb
        d = [np.ascontiguousarray(a, np.float64) for a in (x, y, z, h)]
        npp = np.ascontiguousarray(np_, np.int32)
        cnt = np.array([len(neigh[i]) if i < nc + nv else 0 for i in range(S)], np.int32)
        first = np.concatenate([[0], np.cumsum(cnt)[:-1]]).astype(np.int32)
        flat = np.array([m for i in range(nc + nv) for m in neigh[i]] or [0], np.int32)
        assert self.L.orc_set_comp(self.h, nc, nv, nb, *[_p(a) for a in d], _p(npp, _I),
                                   _p(first, _I), _p(cnt, _I), _p(flat, _I), int(cnt.sum())) == 0

    SEARCH_ERRORS = {-2: "too many neighbours in an octant", -3: "not enough neighbours",
                     -4: "invalid neighbour"}

    def neighbors_status(self):
        """(status, nbr, dist): status 0 or the search's error code."""
        c = self.comp()
        M = c["nc"] + c["nv"]
        lm = self.p.lm
        nbr = np.zeros(M * lm, np.int32)
        dist = np.zeros(M * lm)
        r = self.L.orc_neighbors(self.h, _p(nbr, _I), _p(dist))
        return r, nbr.reshape(M, lm), dist.reshape(M, lm)

    def neighbors(self):
        c = self.comp()
        M = c["nc"] + c["nv"]
        lm = self.p.lm
        nbr = np.zeros(M * lm, np.int32)
        dist = np.zeros(M * lm)
        r = self.L.orc_neighbors(self.h, _p(nbr, _I), _p(dist))
        if r:
            raise RuntimeError("oracle neighbour search failed: %d" % r)
        return nbr.reshape(M, lm), dist.reshape(M, lm)

    def set_constants(self, c_dens, c_pot, c_field):
        self.L.orc_set_constants(self.h, c_dens, c_pot, c_field)

    def build_system(self):
        c = self.comp()
        M = c["nc"] + c["nv"]
        lm = self.p.lm
        cols = np.zeros(M * lm, np.int32)
        a1 = np.zeros(M * lm)
        a2 = np.zeros(M * lm)
        vec = [np.zeros(M) for _ in range(4)]
        r = self.L.orc_build_system(self.h, _p(cols, _I), _p(a1), _p(a2), *[_p(v) for v in vec])
        if r:
            raise RuntimeError("oracle build_system failed: %d" % r)
        return dict(cols=cols.reshape(M, lm), a1=a1.reshape(M, lm), a2=a2.reshape(M, lm),
                    d1=vec[0], d2=vec[1], rhs1=vec[2], pbnd=vec[3])

    def solve(self):
        info = OrcSolveInfo()
        r = self.L.orc_solve(self.h, C.byref(info))
        if r:
            raise RuntimeError("oracle solve failed: %d" % r)
        return info

    def fields(self):
        c = self.comp()
        M = c["nc"] + c["nv"]
        rho, phi, g, H = np.zeros(M), np.zeros(M), np.zeros(3 * M), np.zeros(6 * M)
        assert self.L.orc_comp_fields(self.h, _p(rho), _p(phi), _p(g), _p(H)) == 0
        return dict(rho=rho, phi=phi, g=g.reshape(M, 3), H=H.reshape(M, 6))

    def differentiate(self, phi):
        """Differentiate (solver.cpp:295-362) on a given potential."""
        phi = np.ascontiguousarray(phi, np.float64)
        assert self.L.orc_set_potential(self.h, _p(phi)) == 0
        r = self.L.orc_differentiate(self.h)
        if r:
            raise RuntimeError("oracle differentiate failed: %d" % r)

    def set_gradients(self, g, H):
        g = np.ascontiguousarray(g, np.float64).ravel()
        H = np.ascontiguousarray(H, np.float64).ravel()
        assert self.L.orc_set_gradients(self.h, _p(g), _p(H)) == 0

    def move_velocity(self, cv, dt):
        return self.L.orc_move_velocity(self.h, cv, dt)

    def move_position(self, cx, dt):
        return self.L.orc_move_position(self.h, cx, dt)

    def run(self, start_a, end_a, delta_a, max_steps=0):
        timing = np.zeros(3)
        steps = C.c_int()
        r = self.L.orc_run(self.h, start_a, end_a, delta_a, max_steps, _p(timing), C.byref(steps))
        if r:
            raise RuntimeError("oracle run failed: %d" % r)
        return timing, steps.value


def key64(x, y, z, bbox, depth):
    b = np.ascontiguousarray(bbox, dtype=np.float64)
    return lib().orc_key64(x, y, z, _p(b), depth)


def wlsq(B, b):
    B = np.ascontiguousarray(B, dtype=np.float64).copy()
    b = np.ascontiguousarray(b, dtype=np.float64).copy()
    m, n = B.shape
    nrhs = b.shape[1] if b.ndim == 2 else 1
    info = lib().orc_wlsq(m, n, nrhs, _p(B), _p(b))
    return info, b.reshape(m, nrhs)[:n]


def quicksort(a, b):
    a = np.ascontiguousarray(a, dtype=np.float64).copy()
    b = np.ascontiguousarray(b, dtype=np.int32).copy()
    lib().orc_quicksort(len(a), _p(a), _p(b, _I))
    return a, b


def gmres_csr(rowptr, col, val, rhs, rtol=1e-5, atol=1e-50, dtol=1e5, max_it=10000, restart=30):
    rowptr = np.ascontiguousarray(rowptr, np.int32)
    col = np.ascontiguousarray(col, np.int32)
    val = np.ascontiguousarray(val, np.float64)
    rhs = np.ascontiguousarray(rhs, np.float64)
    x = np.zeros(len(rhs))
    its, reason, rn = C.c_int(), C.c_int(), C.c_double()
    lib().orc_gmres_csr(len(rhs), _p(rowptr, _I), _p(col, _I), _p(val), _p(rhs), _p(x), rtol,
                        atol, dtol, max_it, restart, C.byref(its), C.byref(reason), C.byref(rn))
    return x, its.value, reason.value, rn.value


def coefficients(params, a):
    out = [C.c_double() for _ in range(5)]
    lib().orc_coefficients(C.byref(params), a, *[C.byref(o) for o in out])
    return [o.value for o in out]


# ---------------------------------------------------------------- reference
_ref = None
REF_SO = os.path.join(HERE, "_ref", "libref.so")


def ref_available():
    return os.path.exists(REF_SO)


def ref():
    global _ref
    if _ref is None:
        _ref = C.CDLL(REF_SO)
        R = _ref
        R.ref_tree.argtypes = [C.c_int, _D, _D, _D, _D, C.c_int, _U64, _I, _I, _I, _U64] + \
            [_I] * 5 + [_D, _D, C.c_int]
        R.ref_new.restype = C.c_void_p
        R.ref_new.argtypes = [C.c_int] + [_D] * 7 + [C.c_int, C.c_int]
        R.ref_select.argtypes = [C.c_void_p, C.c_double, C.c_int]
        R.ref_comp_sizes.argtypes = [C.c_void_p, _I, _I, _I, _I]
        R.ref_comp_get.argtypes = [C.c_void_p] + [_D] * 4 + [_I] * 6
        R.ref_move.argtypes = [C.c_void_p, _D, _D, C.c_double, C.c_double, C.c_double, _D, _D]
        R.ref_particles.argtypes = [C.c_void_p] + [_D] * 9
        R.ref_free.argtypes = [C.c_void_p]
        R.ref_view.argtypes = [C.c_char_p, C.c_double, C.c_double, _D, C.c_int] + [_D] * 9 + \
            [C.c_int] + [_D] * 3 + [_I] * 3 + [_D] * 5
    return _ref


def ref_view(out_dir, t0, t1, bbox, pp, cp):
    """The reference's Viewer (viewer.cpp) on the given arrays: pp = 9 arrays
    (x, y, z, v, a), cp = dict(x, y, z, np, type, first, density, potential,
    p_x, p_y, p_z)."""
    R = ref()
    pp = [np.ascontiguousarray(a, np.float64) for a in pp]
    S = len(cp["x"])
    d = {k: np.ascontiguousarray(cp[k], np.float64) for k in
         ("x", "y", "z", "density", "potential", "p_x", "p_y", "p_z")}
    i = {k: np.ascontiguousarray(cp[k], np.int32) for k in ("np", "type", "first")}
    bb = np.ascontiguousarray(bbox, np.float64)
    return R.ref_view(out_dir.encode(), t0, t1, _p(bb), len(pp[0]), *[_p(a) for a in pp], S,
                      _p(d["x"]), _p(d["y"]), _p(d["z"]), _p(i["np"], _I), _p(i["type"], _I),
                      _p(i["first"], _I), _p(d["density"]), _p(d["potential"]), _p(d["p_x"]),
                      _p(d["p_y"]), _p(d["p_z"]))


def ref_tree(x, y, z, bbox, depth, cap=2_000_000):
    n = len(x)
    x, y, z = [np.ascontiguousarray(a, np.float64) for a in (x, y, z)]
    bb = np.ascontiguousarray(bbox, np.float64)
    skey = np.zeros(n, np.uint64)
    sidx = np.zeros(n, np.int32)
    ln = C.c_int()
    nn = np.zeros(depth + 1, np.int32)
    key = np.zeros(cap, np.uint64)
    ints = [np.zeros(cap, np.int32) for _ in range(5)]
    lo = np.zeros(3 * cap)
    hi = np.zeros(3 * cap)
    rc = ref().ref_tree(n, _p(x), _p(y), _p(z), _p(bb), depth, _p(skey, _U64), _p(sidx, _I),
                        C.byref(ln), _p(nn, _I), _p(key, _U64), *[_p(a, _I) for a in ints],
                        _p(lo), _p(hi), cap)
    assert rc == 0
    n2 = ln.value
    return dict(skey=skey, sidx=sidx, numnodes=nn, key=key[:n2], depth=ints[0][:n2],
                first=ints[1][:n2], count=ints[2][:n2], fchild=ints[3][:n2],
                nchild=ints[4][:n2], lo=lo[:3 * n2].reshape(-1, 3), hi=hi[:3 * n2].reshape(-1, 3))


class RefMover:
    """The reference's Mover (and the Octree it owns) over caller particles."""

    def __init__(self, x, y, z, vx, vy, vz, bbox, depth, bc):
        self.n = len(x)
        arrs = [np.ascontiguousarray(a, np.float64) for a in (x, y, z, vx, vy, vz)]
        bb = np.ascontiguousarray(bbox, np.float64)
        self.R = ref()
        self.h = self.R.ref_new(self.n, *[_p(a) for a in arrs], _p(bb), depth, bc)

    def __del__(self):
        if getattr(self, "h", None):
            self.R.ref_free(self.h)
            self.h = None

    def select(self, c, maxnc=10000):
        return self.R.ref_select(self.h, c, maxnc)

    def comp(self):
        nc, nv, nb, nn = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        self.R.ref_comp_sizes(self.h, C.byref(nc), C.byref(nv), C.byref(nb), C.byref(nn))
        S = nc.value + nv.value + nb.value
        d = [np.zeros(S) for _ in range(4)]
        i = [np.zeros(S, np.int32) for _ in range(5)]
        n24 = np.zeros(max(nn.value, 1), np.int32)
        self.R.ref_comp_get(self.h, *[_p(a) for a in d], *[_p(a, _I) for a in i], _p(n24, _I))
        return dict(nc=nc.value, nv=nv.value, nb=nb.value, x=d[0], y=d[1], z=d[2], h=d[3],
                    np=i[0], first=i[1], type=i[2], fneigh=i[3], nneigh=i[4],
                    neigh24=n24[:nn.value])

    def move(self, g, H, cv, cx, dt):
        g = np.ascontiguousarray(g, np.float64)
        H = np.ascontiguousarray(H, np.float64)
        mdv, mdx = C.c_double(), C.c_double()
        self.R.ref_move(self.h, _p(g), _p(H), cv, cx, dt, C.byref(mdv), C.byref(mdx))
        return mdv.value, mdx.value

    def particles(self):
        out = [np.zeros(self.n) for _ in range(9)]
        self.R.ref_particles(self.h, *[_p(a) for a in out])
        return out
