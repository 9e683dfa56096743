"""ctypes wrapper of the CPU oracle (liboracle.so built from stokes_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and by
bench.py's cpu_baseline leg, always as the checker or the timed CPU baseline,
never as part of the product path.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")

_lib = None


def build(force: bool = False) -> str:
    if force or not os.path.exists(LIB) or os.path.getmtime(LIB) < os.path.getmtime(
            os.path.join(HERE, "stokes_oracle.c")):
        subprocess.check_call(["make", "-s", "-C", HERE, "-B" if force else "all"])
    return LIB


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        dp = C.POINTER(C.c_double)
        fp = C.POINTER(C.c_float)
        ip = C.POINTER(C.c_int)
        L.orc_create.restype = C.c_void_p
        L.orc_create.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, dp, dp, dp, C.c_int, C.c_int, C.c_int]
        L.orc_destroy.argtypes = [C.c_void_p]
        L.orc_set_vonly.argtypes = [C.c_void_p, C.c_int]
        L.orc_threads.restype = C.c_int
        L.orc_threads.argtypes = []
        L.orc_set_knobs.argtypes = [C.c_void_p] + [C.c_double] * 7 + [C.c_int] + [C.c_double] * 3
        L.orc_level_info.argtypes = [C.c_void_p, C.c_int, ip]
        L.orc_grid.restype = dp
        L.orc_grid.argtypes = [C.c_void_p, C.c_int, C.c_int]
        L.orc_array.restype = C.c_void_p
        L.orc_array.argtypes = [C.c_void_p, C.c_char_p]
        L.orc_bc_line.argtypes = [C.c_void_p] + [C.c_int] * 7 + [C.c_float, C.c_float] + [C.c_int] * 3
        L.orc_ronurecalc.argtypes = [C.c_void_p, C.c_long, fp, fp, fp, fp, fp, C.c_double]
        L.orc_ronurestrict.argtypes = [C.c_void_p, C.c_int]
        L.orc_assemble.argtypes = [C.c_void_p]
        L.orc_get_row.argtypes = [C.c_void_p, C.c_long, ip, dp, dp]
        L.orc_gausmat4.argtypes = [C.c_void_p, C.c_int, C.c_int]
        L.orc_gausmat4_colored.argtypes = [C.c_void_p, C.c_int, C.c_int]
        L.orc_residual.argtypes = [C.c_void_p, C.c_int]
        L.orc_sweep_tiles.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]
        L.orc_eq_box.argtypes = [C.c_void_p, C.c_int, C.c_void_p]
        L.orc_set_vonly.argtypes = [C.c_void_p, C.c_int]
        L.orc_bc_pass.argtypes = [C.c_void_p, C.c_int, dp]
        L.orc_vprestrict.argtypes = [C.c_void_p, C.c_int]
        L.orc_vpprolong.argtypes = [C.c_void_p, C.c_int, C.c_int]
        L.orc_coarse_factor.argtypes = [C.c_void_p]
        L.orc_coarse_solve.argtypes = [C.c_void_p]
        L.orc_dense_lu.argtypes = [C.c_int, dp, ip]
        L.orc_dense_lu_solve.argtypes = [C.c_int, dp, ip, dp]
        L.orc_vcycle.argtypes = [C.c_void_p, ip, ip, C.c_int, C.c_int, C.c_int, C.c_int]
        L.orc_relres.restype = C.c_double
        L.orc_relres.argtypes = [C.c_void_p, dp]
        L.orc_relres_rows.restype = C.c_double
        L.orc_relres_rows.argtypes = [C.c_void_p, dp, dp]
        L.orc_post.argtypes = [C.c_void_p, dp, dp, dp, dp]
        L.orc_move_markers.argtypes = [C.c_void_p, dp, dp, C.c_long, fp, fp, fp, fp, fp, ip, fp, C.c_double,
                                       C.c_int, C.c_int]
        L.orc_interp.argtypes = [C.c_void_p, dp, dp, C.c_double, C.c_double, C.c_double, dp, dp]
        L.orc_solve.argtypes = [C.c_void_p, ip, ip, C.c_int, C.c_double, C.c_int, dp]
        L.orc_band_solve.argtypes = [C.c_void_p, dp, C.POINTER(C.c_long)]
        L.orc_band_solve_mt.argtypes = [C.c_void_p, dp, C.POINTER(C.c_long), C.c_int]
        L.orc_gcr.argtypes = [C.c_void_p, ip, ip, C.c_int, C.c_double, C.c_double, C.c_int, C.c_int, C.c_int,
                              dp, dp, dp]
        L.orc_hydrostatic.argtypes = [C.c_void_p, dp]
        L.orc_weights.argtypes = [C.c_void_p, dp]
        # heat-equation multigrid (heat_oracle.c)
        L.orh_create.restype = C.c_void_p
        L.orh_create.argtypes = [C.c_void_p, C.c_int]
        L.orh_destroy.argtypes = [C.c_void_p]
        L.orh_set.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_double]
        L.orh_array.restype = C.c_void_p
        L.orh_array.argtypes = [C.c_void_p, C.c_char_p]
        L.orh_nbond.restype = C.c_int
        L.orh_nbond.argtypes = [C.c_void_p]
        L.orh_bc_line.argtypes = [C.c_void_p] + [C.c_int] * 6 + [C.c_float, C.c_float] + [C.c_int] * 3
        L.orh_rocprestrict.argtypes = [C.c_void_p, C.c_int]
        L.orh_row.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_long), dp]
        L.orh_assemble.argtypes = [C.c_void_p]
        L.orh_residual.argtypes = [C.c_void_p, C.c_int]
        L.orh_gausmatt.argtypes = [C.c_void_p, C.c_int, C.c_int]
        L.orh_gausmatt_colored.argtypes = [C.c_void_p, C.c_int, C.c_int]
        L.orh_trestrict.argtypes = [C.c_void_p, C.c_int]
        L.orh_tprolong.argtypes = [C.c_void_p, C.c_int, C.c_int]
        L.orh_err.argtypes = [C.c_void_p, dp, dp]
        L.orh_step.argtypes = [C.c_void_p, ip, ip, C.c_int, C.c_double, C.c_int, C.c_int, dp, dp]
        L.orh_sources.argtypes = [C.c_void_p, dp, dp, dp, fp, fp, C.c_int, C.c_int, C.c_double, C.c_double,
                                  C.c_double]
        L.orh_markers.argtypes = [C.c_void_p, C.c_long, fp, fp, fp, ip, fp, C.c_double, C.c_double, C.c_double,
                                  C.c_double]
        L.orh_titerate.argtypes = [C.c_void_p, C.c_double, C.c_double, ip, ip, C.c_int, C.c_double, C.c_int,
                                   C.c_int, dp]
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


class Oracle:
    """One hierarchy + assembled operator held by the C oracle."""

    def __init__(self, gx, gy, gz, L, periodic=(0, 0, 0)):
        self.L = L
        self._g = [np.ascontiguousarray(a, dtype=np.float64) for a in (gx, gy, gz)]
        self.h = lib().orc_create(len(gx), len(gy), len(gz), L, _dp(self._g[0]), _dp(self._g[1]),
                                  _dp(self._g[2]), *periodic)
        if not self.h:
            raise ValueError("bad hierarchy")
        self.info = []
        for l in range(L + 1):
            buf = np.zeros(6, dtype=np.int32)
            lib().orc_level_info(self.h, l, _ip(buf))
            self.info.append(tuple(int(v) for v in buf))
        self.ntot = self.info[L][5] + self.info[L][4]

    def __del__(self):
        try:
            if getattr(self, "h", None):
                lib().orc_destroy(self.h)
                self.h = None
        except Exception:
            pass

    def knobs(self, p0koef=0.1, p1koef=0.0, p2koef=0.1, v0koef=1.0, v1koef=1.0, nubeg=1.0,
              nuend=1e4, viscmod=0, gravity=(0.0, 0.0, 0.0)):
        lib().orc_set_knobs(self.h, p0koef, p1koef, p2koef, v0koef, v1koef, nubeg, nuend, viscmod,
                            *gravity)

    def grid(self, l, axis):
        n = self.info[l][axis]
        p = lib().orc_grid(self.h, l, axis)
        return np.ctypeslib.as_array(p, shape=(n,)).copy()

    def arr(self, name):
        sizes = {"nu": self.ntot, "ro": self.ntot, "nd": self.ntot, "nxx": self.info[0][4],
                 "nxy": self.info[0][4], "nxz": self.info[0][4], "nyz": self.info[0][4],
                 "bondm": 4 * self.ntot, "sol0": 4 * self.ntot, "sol1": 4 * self.ntot,
                 "fre0": 4 * self.ntot, "fre1": 4 * self.ntot, "val1": 57 * self.ntot,
                 "lin0": 57 * self.ntot, "num0": 4 * self.ntot, "pos0": 4 * self.ntot}
        types = {"bondm": C.c_int, "lin0": C.c_int, "num0": C.c_int, "pos0": C.c_int,
                 "nu": C.c_float, "ro": C.c_float, "nd": C.c_float, "nxx": C.c_float,
                 "nxy": C.c_float, "nxz": C.c_float, "nyz": C.c_float}
        p = lib().orc_array(self.h, name.encode())
        if name in ("bondv1", "bondn1"):
            nb = self.nbond
            if name == "bondv1":
                return np.ctypeslib.as_array(C.cast(p, C.POINTER(C.c_float)), shape=(2 * (nb + 1),))
            return np.ctypeslib.as_array(C.cast(p, C.POINTER(C.c_int)), shape=(nb + 1,))
        t = types.get(name, C.c_double)
        return np.ctypeslib.as_array(C.cast(p, C.POINTER(t)), shape=(sizes[name],))

    @property
    def nbond(self):
        p = lib().orc_array(self.h, b"nbond")
        return C.cast(p, C.POINTER(C.c_int))[0]

    def bc_lines(self, lines):
        """Boundary lines of the Stokes unknowns (T lines, variable 4, belong to Heat)."""
        for ln in lines:
            if int(ln[0]) == 4:
                continue
            lib().orc_bc_line(self.h, *[int(v) for v in ln[:7]], float(ln[7]), float(ln[8]),
                              *[int(v) for v in ln[9:12]])

    def ronurecalc(self, mx, my, mz, mnu, mro, nukoef=0.0):
        a = [np.ascontiguousarray(v, dtype=np.float32) for v in (mx, my, mz, mnu, mro)]
        rc = lib().orc_ronurecalc(self.h, len(a[0]), *[_fp(v) for v in a], nukoef)
        if rc:
            raise RuntimeError("ronurecalc failed")

    def ronurestrict_all(self):
        for l in range(1, self.L + 1):
            lib().orc_ronurestrict(self.h, l)

    def ronurestrict(self, l):
        lib().orc_ronurestrict(self.h, l)

    def assemble(self):
        lib().orc_assemble(self.h)

    def row(self, r):
        idx = np.zeros(64, dtype=np.int32)
        val = np.zeros(64, dtype=np.float64)
        rhs = np.zeros(1, dtype=np.float64)
        n = lib().orc_get_row(self.h, r, _ip(idx), _dp(val), _dp(rhs))
        return idx[:n].copy(), val[:n].copy(), float(rhs[0])

    def gausmat4(self, am, l):
        lib().orc_gausmat4(self.h, am, l)

    def gausmat4_colored(self, am, l):
        """gausmat4 in the device's multicolour order over the row store."""
        rc = lib().orc_gausmat4_colored(self.h, am, l)
        if rc:
            raise RuntimeError("boundary rows form a cycle")

    def residual(self, l):
        lib().orc_residual(self.h, l)

    def sweep_tiles(self, am, l, ty=64, tx=4, tz=4):
        lib().orc_sweep_tiles(self.h, am, l, ty, tx, tz)

    def set_vonly(self, flag):
        """Velocity rows only (pressure rows left out of sweeps and boundary passes)."""
        lib().orc_set_vonly(self.h, int(flag))

    def eq_box(self, l):
        """(y0, x0, z0, y1, x1, z1) of the nodes with equation rows on level l."""
        b = np.zeros(6, np.int32)
        lib().orc_eq_box(self.h, l, b.ctypes.data)
        return tuple(int(v) for v in b)

    def vprestrict(self, l):
        lib().orc_vprestrict(self.h, l)

    def vpprolong(self, l, ynreset=0):
        lib().orc_vpprolong(self.h, l, ynreset)

    def coarse_factor(self):
        return lib().orc_coarse_factor(self.h)

    def coarse_solve(self):
        return lib().orc_coarse_solve(self.h)

    def vcycle(self, pre, post, mode=0, tile=(64, 4, 4)):
        p = np.ascontiguousarray(pre, dtype=np.int32)
        q = np.ascontiguousarray(post, dtype=np.int32)
        lib().orc_vcycle(self.h, _ip(p), _ip(q), mode, *tile)

    def relres(self):
        b = np.zeros(3)
        r = lib().orc_relres(self.h, _dp(b))
        return r, b

    def relres_rows(self, x):
        """Relative residual of finest-level state x with rows built on the
        fly (levels too large for the row store)."""
        x = np.ascontiguousarray(x, dtype=np.float64)
        assert x.size == 4 * self.info[0][4]
        b = np.zeros(3)
        r = lib().orc_relres_rows(self.h, _dp(x), _dp(b))
        return r, b

    def post(self, x, stresses=True):
        """Post-solve pass on level 0 (orc_post): stress and strain-rate planes
        (6, n0) each (xx yy zz xy xz yz; None without ``stresses``) and the 20
        norm sums and extrema."""
        x = np.ascontiguousarray(x, dtype=np.float64)
        n0 = self.info[0][4]
        assert x.size == 4 * n0
        out = np.zeros(20)
        sig = np.zeros((6, n0)) if stresses else None
        eps = np.zeros((6, n0)) if stresses else None
        lib().orc_post(self.h, _dp(x), _dp(sig) if stresses else None, _dp(eps) if stresses else None, _dp(out))
        return sig, eps, out

    def interp(self, x, eps, p):
        """Velocity (3) and strain rates (6) at point p from solution x and the
        strain-rate planes eps (6, n0) (vxyzcalc1 / epscalc1)."""
        x = np.ascontiguousarray(x, dtype=np.float64)
        eps = np.ascontiguousarray(eps, dtype=np.float64)
        v, e = np.zeros(3), np.zeros(6)
        lib().orc_interp(self.h, _dp(x), _dp(eps), float(p[0]), float(p[1]), float(p[2]), _dp(v), _dp(e))
        return v, e

    def move_markers(self, x, eps, mx, my, mz, marke=None, markk=None, markt=None, timestep=1.0, markmod=0,
                     outgrid=0, want_eii=True):
        """One movemark step of the markers (arrays updated in place, float32);
        returns the strain-rate invariant at the start positions."""
        x = np.ascontiguousarray(x, dtype=np.float64)
        eps = np.ascontiguousarray(eps, dtype=np.float64)
        n = len(mx)
        for a in (mx, my, mz, marke, markk):
            assert a is None or (a.dtype == np.float32 and a.flags.c_contiguous and len(a) == n)
        eii = np.zeros(n, np.float32) if want_eii else None
        mt = None if markt is None else np.ascontiguousarray(markt, dtype=np.int32)
        opt = lambda a, f: f(a) if a is not None else None
        self.wraps_out = lib().orc_move_markers(self.h, _dp(x), _dp(eps), n, _fp(mx), _fp(my), _fp(mz),
                                                opt(marke, _fp), opt(markk, _fp), opt(mt, _ip), opt(eii, _fp),
                                                float(timestep), int(markmod), int(outgrid))
        return eii

    def set_fields(self, nu, ro, nxx, nxy, nxz, nyz):
        """Finest-level node properties given directly (instead of ronurecalc)."""
        n0 = self.info[0][4]
        for name, v in (("nu", nu), ("ro", ro), ("nxx", nxx), ("nxy", nxy), ("nxz", nxz), ("nyz", nyz)):
            self.arr(name)[:n0] = np.asarray(v, dtype=np.float32)[:n0]

    def solve(self, pre, post, mode=0, tol=1e-8, maxcyc=200):
        p = np.ascontiguousarray(pre, dtype=np.int32)
        q = np.ascontiguousarray(post, dtype=np.int32)
        hist = np.zeros(maxcyc + 1)
        it = lib().orc_solve(self.h, _ip(p), _ip(q), mode, tol, maxcyc, _dp(hist))
        return it, hist[: abs(it) + 1]

    def gcr(self, pre, post, mode=1, omega=1.0, tol=1e-8, maxit=300, restart=32, hydro=1):
        p = np.ascontiguousarray(pre, dtype=np.int32)
        q = np.ascontiguousarray(post, dtype=np.int32)
        x = np.zeros(4 * self.info[0][4])
        hist = np.zeros(maxit + 1)
        spi = C.c_double()
        it = lib().orc_gcr(self.h, _ip(p), _ip(q), mode, omega, tol, maxit, restart, hydro, _dp(x), _dp(hist),
                           C.byref(spi))
        return it, x, hist[: max(it, 0) + 1], spi.value

    def band_solve(self, threads=None):
        """Banded direct solve (gausmat3): serial, or with the reference's
        thread structure when ``threads`` is given (0: the OpenMP default)."""
        x = np.zeros(4 * self.info[0][4])
        bw = C.c_long(0)
        if threads is None:
            rc = lib().orc_band_solve(self.h, _dp(x), C.byref(bw))
        else:
            rc = lib().orc_band_solve_mt(self.h, _dp(x), C.byref(bw), int(threads))
        return rc, x, bw.value


def dense_lu(a):
    a = np.array(a, dtype=np.float64, order="C")
    n = a.shape[0]
    piv = np.zeros(n, dtype=np.int32)
    rc = lib().orc_dense_lu(n, _dp(a), _ip(piv))
    return rc, a, piv


def dense_lu_solve(lu, piv, b):
    b = np.array(b, dtype=np.float64)
    lib().orc_dense_lu_solve(lu.shape[0], _dp(lu), _ip(piv), _dp(b))
    return b


def from_problem(pb, L=None):
    """Oracle instance for a setups.Problem: hierarchy, knobs, BCs, markers,
    restricted properties, assembled operator and factored coarsest level."""
    L = pb.levels if L is None else L
    o = Oracle(pb.gx, pb.gy, pb.gz, L, periodic=getattr(pb, "periodic", (0, 0, 0)))
    k = pb.knobs
    o.knobs(k["p0koef"], k["p1koef"], k["p2koef"], k["v0koef"], k["v1koef"], k["nubeg"], k["nuend"],
            k["viscmod"], pb.gravity)
    o.bc_lines(pb.bc)
    o.ronurecalc(pb.mx, pb.my, pb.mz, pb.mnu, pb.mro)
    o.ronurestrict_all()
    o.assemble()
    rc = o.coarse_factor()
    if rc:
        raise RuntimeError(f"coarse LU failed rc={rc}")
    return o


def post_norms(out, divvmax=0.0):
    """The reference's norms from orc_post's sums (parallel/move3mg.c:3075-3086,
    3232-3238): maxvxyz = RMS of the three velocity ranges (1e-8 if not
    positive, DIVVMAX if that is positive); stoksum = RMS momentum error,
    stoksum1 = RMS of error*h^2/eta over maxvxyz, stoksum2 = RMS of their
    product, contsum/contsumb = RMS div(v)*h over maxvxyz on continuity cells /
    cells with a pressure boundary equation, bondsum = RMS boundary error."""
    dvx, dvy, dvz = out[11] - out[10], out[13] - out[12], out[15] - out[14]
    mv = ((dvx * dvx + dvy * dvy + dvz * dvz) / 3.0) ** 0.5
    if mv <= 0:
        mv = 1e-8
    if divvmax > 0:
        mv = divvmax
    rms = lambda s, n: (s / n) ** 0.5 if n > 0 else 0.0
    return {"stoksum": rms(out[0], out[3]), "stoksum1": rms(out[1], out[3]) / mv,
            "stoksum2": rms(out[2], out[3]) / mv, "contsum": rms(out[4], out[5]) / mv,
            "contsumb": rms(out[6], out[7]) / mv, "bondsum": rms(out[8], out[9]),
            "stoknum": out[3], "contnum": out[5], "contnumb": out[7], "bondnum": out[9], "maxvxyz": mv,
            "minvx": out[10], "maxvx": out[11], "minvy": out[12], "maxvy": out[13], "minvz": out[14],
            "maxvz": out[15], "minpr": out[16], "maxpr": out[17], "minnu": out[18], "maxnu": out[19]}


class Heat:
    """The heat-equation multigrid of the reference (parallel/heat3mg.c) on
    levels 0..Lt of an Oracle's hierarchy: T boundary lines, FP32 ro/cp/kt
    (all levels), tk/tk0/tk1 (level 0), 7-slot rows, gausmatt, trestrict,
    tprolong, rocprestrict and the titerate step."""

    def __init__(self, oracle: "Oracle", Lt: int):
        self.o = oracle
        self.Lt = Lt
        self.h = lib().orh_create(oracle.h, Lt)
        if not self.h:
            raise ValueError("bad heat levels")
        self.ntot = oracle.info[Lt][5] + oracle.info[Lt][4]
        self.n0 = oracle.info[0][4]

    def __del__(self):
        if getattr(self, "h", None):
            lib().orh_destroy(self.h)
            self.h = None

    def set(self, timestep, t0koef=1.0, t1koef=1.0):
        lib().orh_set(self.h, timestep, t0koef, t1koef)

    def arr(self, name):
        p = lib().orh_array(self.h, name.encode())
        n = self.n0 if name in ("tk", "tk0", "tk1", "tk2") else self.ntot
        if name in ("val1", "lin0"):
            n *= 7
        if name in ("bv1", "bn1"):
            n = (2 if name == "bv1" else 1) * (lib().orh_nbond(self.h) + 1)
        ct = {"tbondm": C.c_int, "bn1": C.c_int, "lin0": C.c_int, "num0": C.c_int, "bv1": C.c_float,
              "ro": C.c_float, "cp": C.c_float, "kt": C.c_float, "tk": C.c_float, "tk0": C.c_float,
              "tk1": C.c_float, "tk2": C.c_float}.get(name, C.c_double)
        return np.ctypeslib.as_array(C.cast(p, C.POINTER(ct)), shape=(n,))

    def bc_lines(self, lines):
        """Boundary lines of variable T (4) in the setups.BcLine form; others are skipped."""
        for ln in lines:
            if int(ln[0]) != 4:
                continue
            lib().orh_bc_line(self.h, *[int(v) for v in ln[1:7]], float(ln[7]), float(ln[8]),
                              *[int(v) for v in ln[9:12]])

    def set_fields(self, ro, cp, kt, tk, tk1):
        """Level-0 properties, temperature at the step start (also tk0) and source term."""
        n0 = self.n0
        for name, v in (("ro", ro), ("cp", cp), ("kt", kt)):
            self.arr(name)[:n0] = np.asarray(v, dtype=np.float32)[:n0]
        self.arr("tk")[:] = np.asarray(tk, dtype=np.float32)[:n0]
        self.arr("tk0")[:] = np.asarray(tk, dtype=np.float32)[:n0]
        self.arr("tk1")[:] = np.asarray(tk1, dtype=np.float32)[:n0]
        self.arr("tk2")[:] = np.asarray(tk, dtype=np.float32)[:n0]

    def rocprestrict(self, l):
        lib().orh_rocprestrict(self.h, l)

    def row(self, l, m1, m2, m3):
        un1 = (C.c_long * 16)()
        ui1 = (C.c_double * 16)()
        lib().orh_row(self.h, l, m1, m2, m3, un1, ui1)
        k = un1[0]
        return np.array(un1[1:k + 1]), np.array(ui1[1:k + 1]), ui1[0]

    def assemble(self):
        lib().orh_assemble(self.h)

    def residual(self, l):
        lib().orh_residual(self.h, l)

    def gausmatt(self, am, l, colored=False):
        if colored:
            if lib().orh_gausmatt_colored(self.h, am, l):
                raise RuntimeError("cyclic T boundary equations")
        else:
            lib().orh_gausmatt(self.h, am, l)

    def trestrict(self, l):
        lib().orh_trestrict(self.h, l)

    def tprolong(self, l, ynreset=0):
        lib().orh_tprolong(self.h, l, ynreset)

    def err(self):
        a, b = C.c_double(), C.c_double()
        lib().orh_err(self.h, C.byref(a), C.byref(b))
        return a.value, b.value

    def step(self, pre, post, multicyct=1, heatmin=1e-4, maxit=200, colored=True):
        p = np.zeros(12, np.int32)
        q = np.zeros(12, np.int32)
        p[:len(pre)] = pre
        q[:len(post)] = post
        a, b = C.c_double(), C.c_double()
        it = lib().orh_step(self.h, _ip(p), _ip(q), multicyct, heatmin, maxit, 1 if colored else 0, C.byref(a),
                            C.byref(b))
        return int(it), a.value, b.value

    def sources(self, x, sig, eps, ht, et, adiabyn, frictyn, gravity=(0.0, 0.0, 0.0)):
        """tk1 = heaterr's right part (radiogenic ht + adiabatic + frictional heating) from a Stokes
        solution x and its stress / strain-rate planes (6, n0); the temperature tk is tk0."""
        x = np.ascontiguousarray(x, dtype=np.float64)
        sig = np.ascontiguousarray(sig, dtype=np.float64)
        eps = np.ascontiguousarray(eps, dtype=np.float64)
        ht = np.ascontiguousarray(ht, dtype=np.float32)
        et = np.ascontiguousarray(et, dtype=np.float32)
        lib().orh_sources(self.h, _dp(x), _dp(sig), _dp(eps), _fp(ht), _fp(et), int(adiabyn), int(frictyn),
                          *[float(v) for v in gravity])

    def markers(self, mx, my, mz, markt, markk, heatdif=1.0, timestep=1.0, zdeep=1e30, tdeep=0.0):
        """Marker temperatures after the step (markk updated in place, float32)."""
        for a in (mx, my, mz, markk):
            assert a.dtype == np.float32 and a.flags.c_contiguous
        mt = np.ascontiguousarray(markt, dtype=np.int32)
        lib().orh_markers(self.h, len(mx), _fp(mx), _fp(my), _fp(mz), _ip(mt), _fp(markk), float(heatdif),
                          float(timestep), float(zdeep), float(tdeep))

    def titerate(self, timestep, maxtkstep, pre, post, multicyct=1, heatmin=1e-4, maxit=200, colored=True):
        """Whole thermal step with the sub-step limit; returns a dict of the step's figures."""
        p = np.zeros(12, np.int32)
        q = np.zeros(12, np.int32)
        p[:len(pre)] = pre
        q[:len(post)] = post
        out = np.zeros(8)
        lib().orh_titerate(self.h, timestep, maxtkstep, _ip(p), _ip(q), multicyct, heatmin, maxit,
                           1 if colored else 0, _dp(out))
        keys = ("substeps", "iterations", "heatsum", "bondsum", "tkstep", "tkmin", "tkmax", "converged")
        return {k: (int(v) if k in ("substeps", "iterations", "converged") else float(v)) for k, v in zip(keys, out)}


def threads() -> int:
    """OpenMP threads of the oracle's parallel loops."""
    return int(lib().orc_threads())
