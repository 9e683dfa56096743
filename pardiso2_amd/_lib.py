"""ctypes binding of libpardiso2_amd.so (the C ABI of include/pardiso2_amd.h).

The library must exist: there is no CPU fallback.  Build it with
``python pardiso2_amd/build.py`` (or ``__graft_entry__.build()``).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "libpardiso2_amd.so")

ERRORS = {0: "ok", -1: "invalid argument", -2: "device out of memory", -3: "HIP error",
          -4: "coarsest matrix singular", -5: "not converged", -6: "NaN/Inf",
          -7: "call order", -8: "no bound context",
          -9: "an outer-layer unknown has no boundary equation (walls on the outer lines)"}

EXPORTS = [
    "stk_create", "stk_destroy", "stk_level_info", "stk_set_knobs", "stk_default_opts", "stk_bc_line",
    "stk_set_bc_tables", "stk_get_bc_tables", "stk_set_fields", "stk_markers_to_nodes",
    "stk_markers_to_nodes_noalloc", "stk_get_fields",
    "stk_prepare", "stk_set_solution", "stk_get_solution", "stk_solve", "stk_relres",
    "stk_pressure_gauge", "stk_hydrostatic_guess", "stk_time_kernel", "stk_vec_set", "stk_vec_get", "stk_apply", "stk_smooth", "stk_smooth_ex", "stk_vrestrict32", "stk_coarse_solve",
    "stk_bind", "matclear", "gausmat", "gausmat2", "pardisomat", "gausmat4", "vprestrict", "vpprolong",
    "ronurestrict", "vpiterate", "stk_kop_apply", "stk_kcycle_apply", "stk_schur_apply",
    "stk_solve_batch", "stk_mem_info", "stk_autotune", "stk_level_tile", "stk_set_level_tile", "stk_tile_chains",
    "stk_norms_eval", "stk_stress_strain", "stk_bind_arrays", "stk_move_markers",
]

# DIVVMIN, STOKSMIN, DIVVMAX, STOKSMAX of the reference's input (parallel/mode.t3c:79-82)
REF_STOP = dict(divvmin=9e-3, stoksmin=5e1, divvmax=0.0, stoksmax=3e-3)


class Knobs(C.Structure):
    _fields_ = [("p0koef", C.c_double), ("p1koef", C.c_double), ("p2koef", C.c_double),
                ("v0koef", C.c_double), ("v1koef", C.c_double), ("nubeg", C.c_double),
                ("nuend", C.c_double), ("viscmod", C.c_int), ("gx", C.c_double), ("gy", C.c_double),
                ("gz", C.c_double), ("divvmin", C.c_double), ("stoksmin", C.c_double), ("divvmax", C.c_double),
                ("stoksmax", C.c_double)]


class Norms(C.Structure):
    _fields_ = [(k, C.c_double) for k in ("stoksum", "stoksum1", "stoksum2", "contsum", "contsumb", "bondsum",
                                          "stoknum", "contnum", "contnumb", "bondnum", "maxvxyz", "minvx",
                                          "maxvx", "minvy", "maxvy", "minvz", "maxvz", "minpr", "maxpr",
                                          "minnu", "maxnu")] + [("converged", C.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


def _ref_fields():
    dp, fp, ip = C.POINTER(C.c_double), C.POINTER(C.c_float), C.POINTER(C.c_int)
    return ([(k, dp) for k in ("sol0", "fre0", "fre1", "val1")] + [(k, ip) for k in ("lin0", "num0", "pos0")] +
            [(k, fp) for k in ("nu", "ro", "nd", "nxx", "nxy", "nxz", "nyz")] +
            [("bondm", ip), ("bondv1", fp), ("bondn1", ip), ("bondnum", C.POINTER(C.c_long)), ("pr", dp)] +
            [(k, fp) for k in ("vx", "vy", "vz", "sxx", "syy", "szz", "sxy", "sxz", "syz", "exx", "eyy", "ezz",
                               "exy", "exz", "eyz")])


class RefArrays(C.Structure):
    """stk_ref_arrays: pointers to a caller's copies of the reference's globals."""
    _fields_ = _ref_fields()


def bind_arrays(arrays):
    """Bind a dict of numpy arrays (names of stk_ref_arrays, reference dtypes:
    float64 for sol0/fre0/fre1/val1/pr, int32 for lin0/num0/pos0/bondm/bondn1,
    float32 for the node properties, bondv1 and the post-solve outputs, int64
    for bondnum) as the caller's reference arrays; None unbinds.  The arrays
    must stay alive while bound."""
    if arrays is None:
        check(lib().stk_bind_arrays(None), "stk_bind_arrays")
        return None
    want = {C.POINTER(C.c_double): np.float64, C.POINTER(C.c_float): np.float32, C.POINTER(C.c_int): np.int32,
            C.POINTER(C.c_long): np.int64}
    ra = RefArrays()
    for name, typ in RefArrays._fields_:
        a = arrays.get(name)
        if a is None:
            continue
        if a.dtype != want[typ] or not a.flags.c_contiguous:
            raise ValueError(f"{name}: need a contiguous {np.dtype(want[typ])} array")
        setattr(ra, name, a.ctypes.data_as(typ))
    check(lib().stk_bind_arrays(C.byref(ra)), "stk_bind_arrays")
    return ra


class Opts(C.Structure):
    _fields_ = [("precision", C.c_int), ("restart", C.c_int), ("pre", C.c_int * 12),
                ("post", C.c_int * 12), ("tile", C.c_int * 3), ("omega", C.c_double),
                ("cycle", C.c_int), ("coarse_viscmod", C.c_int),
                ("tol", C.c_double), ("maxit", C.c_int), ("verbose", C.c_int),
                ("schur", C.c_int), ("ksweeps", C.c_int), ("kscale", C.c_double), ("krylov", C.c_int),
                ("graph", C.c_int), ("history", C.c_int), ("basis", C.c_int)]


class Stats(C.Structure):
    _fields_ = [("iterations", C.c_int), ("status", C.c_int), ("relres", C.c_double),
                ("relres0", C.c_double), ("seconds", C.c_double), ("blocks", C.c_double * 3),
                ("hist", C.c_double * 256), ("norms", Norms)]


class HeatOpts(C.Structure):
    _fields_ = [("levels", C.c_int), ("cycles", C.c_int), ("pre", C.c_int * 12), ("post", C.c_int * 12),
                ("heatmin", C.c_double), ("t0koef", C.c_double), ("t1koef", C.c_double),
                ("maxtkstep", C.c_double), ("maxit", C.c_int)]


class HeatStats(C.Structure):
    _fields_ = [("substeps", C.c_int), ("iterations", C.c_int), ("converged", C.c_int),
                ("heatsum", C.c_double), ("heatnum", C.c_double), ("bondsum", C.c_double),
                ("bondnum", C.c_double), ("tkstep", C.c_double), ("tkmin", C.c_double), ("tkmax", C.c_double),
                ("seconds", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_lib = None


class StkError(RuntimeError):
    pass


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            raise StkError(f"{LIBPATH} is missing: build it with python pardiso2_amd/build.py")
        L = C.CDLL(LIBPATH)
        vp, dp, fp, ip = C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_float), C.POINTER(C.c_int)
        L.stk_create.argtypes = [C.POINTER(vp), C.c_int, C.c_int, C.c_int, C.c_int, dp, dp, dp, C.c_int,
                                 C.c_int, C.c_int]
        L.stk_destroy.argtypes = [vp]
        L.stk_destroy.restype = None
        L.stk_level_info.argtypes = [vp, C.c_int, ip]
        L.stk_set_knobs.argtypes = [vp, C.POINTER(Knobs)]
        L.stk_default_opts.argtypes = [C.POINTER(Opts), C.c_int]
        L.stk_default_opts.restype = None
        L.stk_bc_line.argtypes = [vp] + [C.c_int] * 7 + [C.c_float, C.c_float] + [C.c_int] * 3
        L.stk_set_bc_tables.argtypes = [vp, ip, C.c_int, fp, ip]
        L.stk_get_bc_tables.argtypes = [vp, ip, ip, fp, ip]
        L.stk_set_fields.argtypes = [vp] + [fp] * 6
        L.stk_markers_to_nodes.argtypes = [vp, C.c_long] + [fp] * 5 + [C.c_double]
        L.stk_markers_to_nodes_noalloc.argtypes = [vp, C.c_long] + [fp] * 5 + [C.c_double]
        L.stk_get_fields.argtypes = [vp] + [fp] * 7
        L.stk_prepare.argtypes = [vp]
        L.stk_set_solution.argtypes = [vp, dp]
        L.stk_get_solution.argtypes = [vp, dp]
        L.stk_solve.argtypes = [vp, C.POINTER(Opts), C.POINTER(Stats)]
        L.stk_relres.argtypes = [vp, dp, dp]
        L.stk_pressure_gauge.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_double]
        L.stk_hydrostatic_guess.argtypes = [vp]
        L.stk_time_kernel.argtypes = [vp, C.c_int, C.c_int, C.c_int, dp]
        L.stk_vec_set.argtypes = [vp, C.c_int, C.c_int, dp]
        L.stk_vec_get.argtypes = [vp, C.c_int, C.c_int, dp]
        L.stk_apply.argtypes = [vp, C.c_int, dp, dp]
        L.stk_smooth.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]
        L.stk_smooth_ex.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]
        L.stk_vrestrict32.argtypes = [vp, C.c_int]
        L.stk_coarse_solve.argtypes = [vp]
        L.stk_bind.argtypes = [vp]
        L.stk_kop_apply.argtypes = [vp, C.c_int, dp, dp]
        L.stk_solve_batch.argtypes = [C.c_int, C.POINTER(vp), C.POINTER(Opts), C.c_int, C.POINTER(Stats), C.c_int,
                                      C.c_int]
        L.stk_mem_info.argtypes = [C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]
        L.stk_autotune.argtypes = [vp, C.c_int, C.c_int]
        L.stk_level_tile.argtypes = [vp, C.c_int, C.c_int, ip]
        L.stk_set_level_tile.argtypes = [vp, C.c_int, C.c_int, ip]
        L.stk_tile_chains.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_int]
        L.stk_kcycle_apply.argtypes = [vp, C.c_int, C.c_int, C.c_double, dp, dp]
        L.stk_schur_apply.argtypes = [vp, C.POINTER(Opts), dp, dp]
        L.stk_norms_eval.argtypes = [vp, C.POINTER(Norms)]
        L.stk_bind_arrays.argtypes = [C.POINTER(RefArrays)]
        L.stk_move_markers.argtypes = [vp, C.c_long, fp, fp, fp, fp, fp, ip, fp, C.c_double, C.c_int, C.c_int]
        L.stk_stress_strain.argtypes = [vp, dp, dp]
        L.stk_heat_default_opts.argtypes = [C.POINTER(HeatOpts)]
        L.stk_heat_default_opts.restype = None
        L.stk_heat_set_fields.argtypes = [vp] + [fp] * 5
        L.stk_heat_get_fields.argtypes = [vp] + [fp] * 6
        L.stk_heat_sources.argtypes = [vp, fp, C.c_int, C.c_int]
        L.stk_heat_markers.argtypes = [vp, C.c_long, fp, fp, fp, ip, fp, C.c_double, C.c_double, C.c_double,
                                       C.c_double]
        L.stk_heat_set_bc_tables.argtypes = [vp, ip, C.c_int, fp, ip]
        L.stk_heat_titerate.argtypes = [vp, C.c_double, C.POINTER(HeatOpts), C.POINTER(HeatStats)]
        L.stk_heat_prepare.argtypes = [vp, C.c_int, C.c_double, C.c_double, C.c_double]
        L.stk_heat_gausmatt.argtypes = [vp, C.c_int, C.c_int]
        L.stk_heat_trestrict.argtypes = [vp, C.c_int]
        L.stk_heat_tprolong.argtypes = [vp, C.c_int, C.c_int]
        L.stk_heat_err.argtypes = [vp, dp, dp]
        L.stk_heat_vec.argtypes = [vp, C.c_int, C.c_int, dp, C.c_int]
        L.stk_heat_time.argtypes = [vp, C.c_int, C.c_int, C.c_int, dp]
        L.matclear.argtypes = [C.c_int, dp]
        L.matclear.restype = None
        L.gausmat.argtypes = [C.c_int, dp, dp]
        L.gausmat.restype = None
        L.gausmat2.argtypes = [C.c_int, C.c_long, C.c_long, C.POINTER(C.c_long), dp]
        L.pardisomat.argtypes = [C.c_int, C.c_long, C.c_long, C.c_long, C.c_int, C.POINTER(C.c_long), dp]
        L.gausmat4.argtypes = [C.c_int, C.c_long, C.c_long, C.c_int]
        for nm in ("vprestrict", "ronurestrict"):
            getattr(L, nm).argtypes = [C.c_int]
            getattr(L, nm).restype = None
        L.vpprolong.argtypes = [C.c_int, C.c_int]
        L.vpprolong.restype = None
        L.vpiterate.argtypes = [C.c_int, C.c_int]
        L.vpiterate.restype = None
        _lib = L
    return _lib


def check(rc, what=""):
    if rc != 0:
        raise StkError(f"{what}: {ERRORS.get(rc, rc)} ({rc})")
    return rc


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def default_opts(levels: int, **kw) -> Opts:
    o = Opts()
    lib().stk_default_opts(C.byref(o), levels)
    for k, v in kw.items():
        if k in ("pre", "post"):
            arr = getattr(o, k)
            for i, x in enumerate(v):
                arr[i] = int(x)
        elif k == "tile":
            for i, x in enumerate(v):
                o.tile[i] = int(x)
        else:
            setattr(o, k, v)
    return o


def heat_default_opts(**kw) -> HeatOpts:
    o = HeatOpts()
    lib().stk_heat_default_opts(C.byref(o))
    for k, v in kw.items():
        if k in ("pre", "post"):
            arr = getattr(o, k)
            for i in range(12):
                arr[i] = 0
            for i, x in enumerate(v):
                arr[i] = int(x)
        else:
            setattr(o, k, v)
    return o


class Stokes:
    """One Stokes problem resident on the GPU (a stk_ctx)."""

    def __init__(self, gx, gy, gz, levels: int, periodic=(0, 0, 0)):
        self.levels = levels
        self.periodic = tuple(int(bool(v)) for v in periodic)
        self._g = [np.ascontiguousarray(a, dtype=np.float64) for a in (gx, gy, gz)]
        h = C.c_void_p()
        check(lib().stk_create(C.byref(h), len(gx), len(gy), len(gz), levels, _dp(self._g[0]),
                               _dp(self._g[1]), _dp(self._g[2]), *self.periodic), "stk_create")
        self.h = h
        self.info = []
        for l in range(levels + 1):
            b = np.zeros(6, dtype=np.int32)
            check(lib().stk_level_info(self.h, l, _ip(b)), "stk_level_info")
            self.info.append(tuple(int(v) for v in b))
        self.ntot = self.info[-1][5] + self.info[-1][4]
        self.n0 = self.info[0][4]

    def close(self):
        if getattr(self, "h", None):
            lib().stk_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def knobs(self, p0koef=0.1, p1koef=0.0, p2koef=0.1, v0koef=1.0, v1koef=1.0, nubeg=1.0, nuend=1e4,
              viscmod=0, gravity=(0.0, 0.0, 0.0), divvmin=REF_STOP["divvmin"], stoksmin=REF_STOP["stoksmin"],
              divvmax=REF_STOP["divvmax"], stoksmax=REF_STOP["stoksmax"]):
        k = Knobs(p0koef, p1koef, p2koef, v0koef, v1koef, nubeg, nuend, viscmod, *gravity, divvmin, stoksmin,
                  divvmax, stoksmax)
        check(lib().stk_set_knobs(self.h, C.byref(k)), "stk_set_knobs")

    def bc_lines(self, lines):
        for ln in lines:
            check(lib().stk_bc_line(self.h, *[int(v) for v in ln[:7]], float(ln[7]), float(ln[8]),
                                    *[int(v) for v in ln[9:12]]), "stk_bc_line")

    def set_fields(self, nu, ro, nxx, nxy, nxz, nyz):
        a = [np.ascontiguousarray(v[: self.n0], dtype=np.float32) for v in (nu, ro, nxx, nxy, nxz, nyz)]
        check(lib().stk_set_fields(self.h, *[_fp(v) for v in a]), "stk_set_fields")

    def markers_to_nodes(self, mx, my, mz, mnu, mro, nukoef=0.0):
        a = [np.ascontiguousarray(v, dtype=np.float32) for v in (mx, my, mz, mnu, mro)]
        check(lib().stk_markers_to_nodes(self.h, len(a[0]), *[_fp(v) for v in a], nukoef),
              "stk_markers_to_nodes")

    def markers_to_nodes_noalloc(self, mx, my, mz, mnu, mro, nukoef=0.0):
        """The same averaging without device allocation (static per-GPU scratch);
        returns the library's status code."""
        a = [np.ascontiguousarray(v, dtype=np.float32) for v in (mx, my, mz, mnu, mro)]
        return lib().stk_markers_to_nodes_noalloc(self.h, len(a[0]), *[_fp(v) for v in a], nukoef)

    def fields(self):
        nu = np.zeros(self.ntot, np.float32)
        ro = np.zeros(self.ntot, np.float32)
        nd = np.zeros(self.ntot, np.float32)
        f = [np.zeros(self.n0, np.float32) for _ in range(4)]
        check(lib().stk_get_fields(self.h, _fp(nu), _fp(ro), _fp(nd), *[_fp(v) for v in f]), "stk_get_fields")
        return dict(nu=nu, ro=ro, nd=nd, nxx=f[0], nxy=f[1], nxz=f[2], nyz=f[3])

    def prepare(self):
        check(lib().stk_prepare(self.h), "stk_prepare")

    def set_solution(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        check(lib().stk_set_solution(self.h, _dp(x)), "stk_set_solution")

    def solution(self):
        x = np.zeros(4 * self.n0)
        check(lib().stk_get_solution(self.h, _dp(x)), "stk_get_solution")
        return x

    def solve(self, opts: Opts = None, raise_on_fail=False, **kw):
        o = opts if opts is not None else default_opts(self.levels, **kw)
        s = Stats()
        rc = lib().stk_solve(self.h, C.byref(o), C.byref(s))
        if raise_on_fail:
            check(rc, "stk_solve")
        elif rc not in (0, -5, -6):
            check(rc, "stk_solve")
        return s

    def relres(self):
        r = C.c_double()
        b = np.zeros(3)
        check(lib().stk_relres(self.h, C.byref(r), _dp(b)), "stk_relres")
        return r.value, b

    def norms(self):
        """The reference's convergence norms of the current solution (stk_norms_eval)."""
        nm = Norms()
        check(lib().stk_norms_eval(self.h, C.byref(nm)), "stk_norms_eval")
        return nm.as_dict()

    def stress_strain(self):
        """Stress and strain-rate planes (6, n0) each: xx yy zz xy xz yz."""
        sig = np.zeros((6, self.n0))
        eps = np.zeros((6, self.n0))
        check(lib().stk_stress_strain(self.h, _dp(sig), _dp(eps)), "stk_stress_strain")
        return sig, eps

    def move_markers(self, mx, my, mz, marke=None, markk=None, markt=None, timestep=1.0, markmod=0, outgrid=0,
                     want_eii=True, eii=None):
        """One marker-advection step through the current solution
        (stk_move_markers): float32 arrays updated in place; returns the
        strain-rate invariant at the start positions (into eii when given:
        entries of markers the step skips keep their values)."""
        n = len(mx)
        for a in (mx, my, mz, marke, markk, eii):
            if a is not None and (a.dtype != np.float32 or not a.flags.c_contiguous or len(a) != n):
                raise ValueError("marker arrays: contiguous float32 of equal length")
        if eii is None:
            eii = np.zeros(n, np.float32) if want_eii else None
        mt = None if markt is None else np.ascontiguousarray(markt, dtype=np.int32)
        opt = lambda a, f: f(a) if a is not None else None
        check(lib().stk_move_markers(self.h, n, _fp(mx), _fp(my), _fp(mz), opt(marke, _fp), opt(markk, _fp),
                                     opt(mt, _ip), opt(eii, _fp), float(timestep), int(markmod), int(outgrid)),
              "stk_move_markers")
        return eii

    def pressure_gauge(self, px, py, pz, pinit=0.0):
        check(lib().stk_pressure_gauge(self.h, px, py, pz, pinit), "stk_pressure_gauge")

    def hydrostatic_guess(self):
        check(lib().stk_hydrostatic_guess(self.h), "stk_hydrostatic_guess")

    def time_kernel(self, which, level=0, reps=20):
        ms = C.c_double()
        check(lib().stk_time_kernel(self.h, which, level, reps, C.byref(ms)), "stk_time_kernel")
        return ms.value

    def vec_set(self, which, level, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        assert x.size == 4 * self.info[level][4]
        check(lib().stk_vec_set(self.h, which, level, _dp(x)), "stk_vec_set")

    def vec_get(self, which, level):
        x = np.zeros(4 * self.info[level][4])
        check(lib().stk_vec_get(self.h, which, level, _dp(x)), "stk_vec_get")
        return x

    def apply(self, level, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.zeros_like(x)
        check(lib().stk_apply(self.h, level, _dp(x), _dp(y)), "stk_apply")
        return y

    def smooth(self, level, sweeps, tile=(32, 4, 4), precision=1, velocity_only=False):
        if precision == 1 and not velocity_only:
            check(lib().stk_smooth(self.h, level, sweeps, *tile), "stk_smooth")
        else:
            check(lib().stk_smooth_ex(self.h, level, sweeps, *tile, precision, int(velocity_only)), "stk_smooth_ex")

    def vrestrict32(self, level):
        check(lib().stk_vrestrict32(self.h, level), "stk_vrestrict32")

    def coarse_solve(self):
        check(lib().stk_coarse_solve(self.h), "stk_coarse_solve")

    def bind(self):
        check(lib().stk_bind(self.h), "stk_bind")

    def autotune(self, precision=0, reps=5):
        check(lib().stk_autotune(self.h, precision, reps), "stk_autotune")
        out = []
        for l in range(self.levels + 1):
            t = np.zeros(4, np.int32)
            check(lib().stk_level_tile(self.h, l, precision, _ip(t)), "stk_level_tile")
            out.append(tuple(int(v) for v in t))
        return out

    def set_tile(self, level, tile, precision=0):
        """Smoother tile (y, x, z, viscosities staged) of a level, by hand."""
        t = np.ascontiguousarray(list(tile) + [0] * (4 - len(tile)), dtype=np.int32)
        check(lib().stk_set_level_tile(self.h, level, precision, _ip(t)), "stk_set_level_tile")

    def tile_chains(self, level, tile):
        """True when sweeps of this tile shape set the level's boundary rows in the tiles."""
        rc = lib().stk_tile_chains(self.h, level, tile[0], tile[1], abs(tile[2]))
        check(min(rc, 0), "stk_tile_chains")
        return rc == 1

    def kop_apply(self, level, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        assert x.size == self.info[level][4]
        y = np.zeros_like(x)
        check(lib().stk_kop_apply(self.h, level, _dp(x), _dp(y)), "stk_kop_apply")
        return y

    def kcycle_apply(self, b, precision=1, ksweeps=2, kscale=2.0):
        b = np.ascontiguousarray(b, dtype=np.float64)
        assert b.size == self.n0
        x = np.zeros_like(b)
        check(lib().stk_kcycle_apply(self.h, precision, ksweeps, kscale, _dp(b), _dp(x)), "stk_kcycle_apply")
        return x

    def schur_apply(self, r, opts=None, **kw):
        o = opts if opts is not None else default_opts(self.levels, **kw)
        r = np.ascontiguousarray(r, dtype=np.float64)
        z = np.zeros_like(r)
        check(lib().stk_schur_apply(self.h, C.byref(o), _dp(r), _dp(z)), "stk_schur_apply")
        return z

    # ---- temperature step (parallel/heat3mg.c) ----
    def heat_set_fields(self, ro, cp, kt, tk, ht):
        a = [np.ascontiguousarray(v, dtype=np.float32)[: self.n0] for v in (ro, cp, kt, tk, ht)]
        check(lib().stk_heat_set_fields(self.h, *[_fp(v) for v in a]), "stk_heat_set_fields")

    def heat_set_bc_tables(self, tbondm, nbond, bondv1, bondn1):
        """T boundary rows from the reference's tables (tbondm = bondm + nodenum8, ntot entries)."""
        tb = np.ascontiguousarray(tbondm, dtype=np.int32)
        assert tb.size == self.ntot
        v1 = np.ascontiguousarray(bondv1, dtype=np.float32)
        n1 = np.ascontiguousarray(bondn1, dtype=np.int32)
        check(lib().stk_heat_set_bc_tables(self.h, _ip(tb), int(nbond), _fp(v1), _ip(n1)), "stk_heat_set_bc_tables")

    def heat_fields(self):
        """ro, cp, kt of every level; tk, tk0, tk1 of level 0."""
        out = [np.zeros(self.ntot, np.float32) for _ in range(3)] + [np.zeros(self.n0, np.float32) for _ in range(3)]
        check(lib().stk_heat_get_fields(self.h, *[_fp(v) for v in out]), "stk_heat_get_fields")
        return dict(zip(("ro", "cp", "kt", "tk", "tk0", "tk1"), out))

    def heat_sources(self, et=None, adiabyn=0, frictyn=0):
        """tk1 = ht + heaterr's adiabatic and frictional terms from the current Stokes solution."""
        e = None if et is None else np.ascontiguousarray(et, dtype=np.float32)[: self.n0]
        check(lib().stk_heat_sources(self.h, _fp(e) if e is not None else None, int(adiabyn), int(frictyn)),
              "stk_heat_sources")

    def heat_markers(self, mx, my, mz, markt, markk, heatdif=1.0, timestep=1.0, zdeep=1e30, tdeep=0.0):
        """Marker temperatures after the step (markk, float32, updated in place)."""
        n = len(mx)
        for a in (mx, my, mz, markk):
            if a.dtype != np.float32 or not a.flags.c_contiguous or len(a) != n:
                raise ValueError("marker arrays: contiguous float32 of equal length")
        mt = None if markt is None else np.ascontiguousarray(markt, dtype=np.int32)
        check(lib().stk_heat_markers(self.h, n, _fp(mx), _fp(my), _fp(mz), _ip(mt) if mt is not None else None,
                                     _fp(markk), float(heatdif), float(timestep), float(zdeep), float(tdeep)),
              "stk_heat_markers")

    def heat_titerate(self, timestep, opts: HeatOpts = None, raise_on_fail=True, **kw):
        o = opts if opts is not None else heat_default_opts(**kw)
        st = HeatStats()
        rc = lib().stk_heat_titerate(self.h, timestep, C.byref(o), C.byref(st))
        if rc and (raise_on_fail or rc != -5):
            check(rc, "stk_heat_titerate")
        return st

    def heat_prepare(self, levels, timestep, t0koef=1.0, t1koef=1.0):
        check(lib().stk_heat_prepare(self.h, levels, timestep, t0koef, t1koef), "stk_heat_prepare")

    def heat_gausmatt(self, level, sweeps):
        check(lib().stk_heat_gausmatt(self.h, level, sweeps), "stk_heat_gausmatt")

    def heat_trestrict(self, level):
        check(lib().stk_heat_trestrict(self.h, level), "stk_heat_trestrict")

    def heat_tprolong(self, level, ynreset=0):
        check(lib().stk_heat_tprolong(self.h, level, ynreset), "stk_heat_tprolong")

    def heat_err(self):
        a, b = C.c_double(), C.c_double()
        check(lib().stk_heat_err(self.h, C.byref(a), C.byref(b)), "stk_heat_err")
        return a.value, b.value

    def heat_vec_get(self, which, level):
        out = np.zeros(self.info[level][4])
        check(lib().stk_heat_vec(self.h, which, level, _dp(out), 0), "stk_heat_vec")
        return out

    def heat_time(self, which, level=0, reps=20):
        ms = C.c_double()
        check(lib().stk_heat_time(self.h, which, level, reps, C.byref(ms)), "stk_heat_time")
        return ms.value

    def heat_vec_set(self, which, level, x):
        a = np.ascontiguousarray(x, dtype=np.float64)
        assert a.size == self.info[level][4]
        check(lib().stk_heat_vec(self.h, which, level, _dp(a), 1), "stk_heat_vec")


SOL0, FRE1, FRE0 = 0, 1, 2


def from_heat_problem(hp):
    """Context holding a setups.HeatProblem: hierarchy, T boundary lines, level-0 fields."""
    s = Stokes(hp.gx, hp.gy, hp.gz, hp.levels, periodic=hp.periodic)
    s.bc_lines(hp.bc)
    s.heat_set_fields(hp.ro, hp.cp, hp.kt, hp.tk, hp.tk1)
    return s


def solve_batch(contexts, opts=None, threads=16, prepare=False):
    """Solve contexts concurrently (stk_solve_batch); opts is one Opts for all
    or a list; prepare=True also prepares each context and seeds the
    hydrostatic guess inside the workers; returns (rc, [Stats])."""
    n = len(contexts)
    hs = (C.c_void_p * n)(*[c.h.value if isinstance(c.h, C.c_void_p) else c.h for c in contexts])
    if opts is None:
        opts = default_opts(contexts[0].levels)
    if isinstance(opts, Opts):
        oa, no = (Opts * 1)(opts), 1
    else:
        oa, no = (Opts * n)(*opts), n
    st = (Stats * n)()
    rc = lib().stk_solve_batch(n, hs, oa, no, st, threads, 1 if prepare else 0)
    if rc not in (0, -5, -6):
        check(rc, "stk_solve_batch")
    return rc, list(st)


def mem_info():
    f, t = C.c_size_t(), C.c_size_t()
    check(lib().stk_mem_info(C.byref(f), C.byref(t)), "stk_mem_info")
    return f.value, t.value


def from_problem(pb, markers=True, knobs=None):
    """Stokes context for a setups.Problem: hierarchy, knobs, boundary lines,
    node fields from the markers (on the device), prepared operators."""
    s = Stokes(pb.gx, pb.gy, pb.gz, pb.levels, periodic=getattr(pb, "periodic", (0, 0, 0)))
    k = dict(pb.knobs)
    if knobs:
        k.update(knobs)
    s.knobs(k["p0koef"], k["p1koef"], k["p2koef"], k["v0koef"], k["v1koef"], k["nubeg"], k["nuend"],
            k["viscmod"], pb.gravity, **{q: k.get(q, v) for q, v in REF_STOP.items()})
    s.bc_lines(pb.bc)
    if markers:
        s.markers_to_nodes(pb.mx, pb.my, pb.mz, pb.mnu, pb.mro)
    s.prepare()
    return s
