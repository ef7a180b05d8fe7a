"""MI355X-native explicit finite-volume Euler solver (the hot path of PSST,
nkmashaev/SupersonicFlowSolver) behind a C-ABI, with a thin Python mirror of
the reference's host flow:

    case = Case.open("wedge80_240.msh", "wedge")   # reader + settings files
    solver = Solver(case)                          # psstGridFV::solver_parameters
    res = solver.run(3000)                         # explicit_solve's loop, on the GPU
    prim = solver.state()                          # u..., p, T, rho per cell

``Case`` wraps the host C++ (Fluent reader, grid flattening, `.inp/.prop/
.par/.bvp/.btq/.sav` parsing); ``Solver`` wraps the device context
(``psst_ctx_*``).  Errors raise ``PsstError`` (the reference throws C++
exceptions at the same points).
"""
import atexit
import ctypes as C
import os
import weakref

import numpy as np

from ._lib import BC, MeshDesc, PsstError, Settings, WedgeParams, lib

__all__ = ["Case", "Solver", "Batch", "PsstError", "WedgeParams", "wedge_params", "write_wedge_msh",
           "write_case_files", "autotune"]


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


# Objects holding library handles that are still open.  At interpreter exit
# they are closed, device objects first, while the HIP runtime is still whole:
# Python's own finalisation may collect them in any order, or not at all,
# and the runtime's exit handlers run after it.
_live = weakref.WeakSet()


def _track(obj):
    _live.add(obj)
    return obj


@atexit.register
def _close_live():
    objs = list(_live)
    objs.sort(key=lambda o: isinstance(o, Case))  # device contexts and batches before host cases
    for o in objs:
        try:
            o.close()
        except Exception:
            pass


def _check(st, ctx=None):
    if st != 0:
        L = lib()
        msg = L.psst_last_error(ctx).decode() if ctx is not None else L.psst_case_last_error().decode()
        if ctx is None and not msg:
            msg = L.psst_last_error(None).decode()
        raise PsstError(st, msg)


def wedge_params(nx=120, ny=40, nz=0, length=3.0, height=1.0, depth=1.0, ramp_start=0.5, ramp_length=0.6,
                 ramp_deg=15.0, jitter=0.0, jitter_seed=1, numbering=0, perm_seed=2, tets=0, periodic_z=0):
    """Parameters of the generated compression-ramp channel (see psst_wedge_params)."""
    return WedgeParams(nx, ny, nz, length, height, depth, ramp_start, ramp_length, ramp_deg, jitter, jitter_seed,
                       numbering, perm_seed, tets, periodic_z)


def write_wedge_msh(params, path):
    """Writes a generated wedge mesh as a Fluent ASCII file the reference reader accepts."""
    _check(lib().psst_wedge_write_msh(C.byref(params), str(path).encode()))


def write_case_files(prefix, *, cfl=0.8, time_mode=2, iterations=2000, save_step=5000, rk_levels=4,
                     rk_custom=None, flux=3, order=2, grad=1, limiter=2, mach=2.5, p=101325.0, T=300.0,
                     molar_mass=0.029, cp=None, init_mode=0, bottom_bc="15", wall_names=("symm", "wedge", "wall", "up"),
                     dim=2, in_name="in", out_name="out", out_p=None, wall_T=None):
    """Writes the five settings files of a wedge case in the reference's formats
    (psstFiniteVolumeGrid.cpp:36-194, :403-717).  gamma = 1.4 is realised through
    cp = 3.5 R/M with the reference's R = 8.31 (quirk Q10); the inflow speed is
    mach * sqrt(gamma rm T) along x (dim velocity components in 3D).
    .bvp/.btq are written without a trailing newline (the reference re-reads
    the last line otherwise, quirk Q11).  The outlet is supersonic (26), or
    subsonic with fixed pressure out_p (25); the walls are adiabatic (18), or
    held at temperature wall_T (11)."""
    R = 8.31
    rm = R / molar_mass
    if cp is None:
        cp = 3.5 * rm
    gamma = cp / (cp - rm)
    u = mach * np.sqrt(gamma * rm * T)
    with open(prefix + ".inp", "w") as f:
        f.write("%d\t\t//0 - initField 1 - savedField\n" % init_mode)
        f.write("%r\t\t//CFL\n" % cfl)
        f.write("%d\t\t//0 - defaultTime; 1 - localTime1; 2 - localTime2;\n" % time_mode)
        f.write("%d\t\t//maxIter\n" % iterations)
        f.write("%d\t\t//iterSaveStep\n" % save_step)
        f.write("%d\t\t//RK level\n" % rk_levels)
        if rk_custom is None:
            f.write("0 \t\t//0 - Standart RK Coeff; 1-Custom RK Coeff\n")
            f.write("0.1 0.2 0.5 1.0 // Coeff for Custom RK (if custom)\n")
        else:
            f.write("1 \t\t//0 - Standart RK Coeff; 1-Custom RK Coeff\n")
            f.write(" ".join(repr(float(a)) for a in rk_custom) + " // Coeff for Custom RK (if custom)\n")
        f.write("%d\t\t//mode: 1-AUSM, 2-HLL, 3-HLLC\n" % flux)
        f.write("%d\t\t//order: 1 - firstOrder, 2 - secondOrder\n" % order)
        f.write("%d\t\t//grad: 1 - GreenGauss, 2 - LeastSquare\n" % grad)
        f.write("%d\t\t//limiters: 1 - VanAldaba, 2 - MinMod, 3 - VanLeer" % limiter)
    zeros = "\t0.0" * (dim - 1)
    lines = ["%s\t\t23\t%r\t%s\t%r\t//u,v,p" % (in_name, float(u), zeros, float(p))]
    lines += ["%s\t\t%s" % (n, bottom_bc) for n in wall_names]
    lines += ["%s\t\t26" % out_name if out_p is None else "%s\t\t25\t%r\t//p" % (out_name, float(out_p))]
    with open(prefix + ".bvp", "w") as f:
        f.write("\n".join(lines))
    walls = ["%s\t\t18" % n if wall_T is None else "%s\t\t11\t%r\t//T" % (n, float(wall_T)) for n in wall_names]
    lines = ["%s\t\t27\t%r\t//T" % (in_name, float(T))] + walls + ["%s\t\t25" % out_name]
    with open(prefix + ".btq", "w") as f:
        f.write("\n".join(lines))
    with open(prefix + ".par", "w") as f:
        f.write("%r%s  //velocity\n%r\t\t//Temperature\n%r\t//pressure\n" % (float(u), " 0.0" * (dim - 1), float(T),
                                                                                float(p)))
    with open(prefix + ".prop", "w") as f:
        f.write("%r\t\t//M\n%r       //cp" % (float(molar_mass), float(cp)))
    return {"u_inf": float(u), "gamma": float(gamma), "rm": float(rm), "cp": float(cp)}


class Case:
    """A mesh plus the reference's settings files, flattened by the host C++."""

    def __init__(self, handle):
        self._h = C.c_void_p(handle)
        _track(self)
        info = (C.c_int64 * 8)()
        _check(lib().psst_case_info(self._h, info))
        self.dim, self.n_cells, self.n_ghosts, self.n_faces = (int(info[i]) for i in range(4))
        self.rk_levels, self.first_iter, self.last_iter, self.save_step = (int(info[i]) for i in range(4, 8))

    @classmethod
    def open(cls, mesh_path, case_prefix):
        h = C.c_void_p()
        _check(lib().psst_case_open(str(mesh_path).encode(), str(case_prefix).encode(), C.byref(h)))
        return cls(h.value)

    @classmethod
    def wedge(cls, params, case_prefix):
        h = C.c_void_p()
        _check(lib().psst_case_open_wedge(C.byref(params), str(case_prefix).encode(), C.byref(h)))
        return cls(h.value)

    def close(self):
        if self._h:
            lib().psst_case_close(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def n_total(self):
        return self.n_cells + self.n_ghosts

    def settings(self):
        s = Settings()
        _check(lib().psst_case_settings(self._h, C.byref(s)))
        return s

    def set_settings(self, s):
        _check(lib().psst_case_set_settings(self._h, C.byref(s)))

    def mesh(self):
        m = MeshDesc()
        _check(lib().psst_case_mesh(self._h, C.byref(m)))
        return m

    def initial_state(self):
        prim = np.zeros((self.dim + 3, self.n_total), dtype=np.float64)
        _check(lib().psst_case_initial_state(self._h, _dp(prim)))
        return prim

    def write_sav(self, prim, iter_value, path):
        prim = np.ascontiguousarray(prim, dtype=np.float64)
        _check(lib().psst_case_write_sav(self._h, _dp(prim), int(iter_value), str(path).encode()))

    def write_dat(self, prim, path):
        """The reference's `.dat` field file (p, T, rho, velocity of every interior and ghost cell)."""
        prim = np.ascontiguousarray(prim, dtype=np.float64)
        assert prim.shape == (self.dim + 3, self.n_total)
        _check(lib().psst_case_write_dat(self._h, _dp(prim), str(path).encode()))

    def write_cas(self, path):
        """The reference's `.cas` file of the mesh (zones, faces, cells, trees of an adapted mesh)."""
        _check(lib().psst_case_write_cas(self._h, str(path).encode()))

    def raw(self):
        """The mesh as read: vertices, faces, cells and zones (file order)."""
        L = lib()
        cnt = (C.c_int64 * 4)()
        _check(L.psst_case_raw_counts(self._h, cnt))
        nv, nf, nc, nz = (int(x) for x in cnt)
        xyz = np.zeros((nv, 3))
        face_nv = np.zeros(nf, dtype=np.int8)
        face_v = np.zeros((nf, 4), dtype=np.int64)
        face_c = np.zeros((nf, 2), dtype=np.int64)
        face_zone = np.zeros(nf, dtype=np.int64)
        cell_zone = np.zeros(nc, dtype=np.int64)
        zid = np.zeros(nz, dtype=np.int64)
        zr = np.zeros((nz, 2), dtype=np.int64)
        zbc = np.zeros(nz, dtype=np.int64)
        names = C.create_string_buffer(65536)
        p64 = lambda a: a.ctypes.data_as(C.POINTER(C.c_int64))
        _check(L.psst_case_raw(self._h, _dp(xyz), face_nv.ctypes.data_as(C.POINTER(C.c_int8)), p64(face_v),
                               p64(face_c), p64(face_zone), p64(cell_zone), p64(zid), p64(zr), p64(zbc), names,
                               65536))
        return dict(xyz=xyz, face_nv=face_nv, face_v=face_v, face_c=face_c, face_zone=face_zone,
                    cell_zone=cell_zone, zone_id=zid, zone_range=zr, zone_bc=zbc,
                    zone_names=names.value.decode().split("\n")[:nz])

    def solve(self):
        """Runs the case like the reference's main(): writes <prefix>_residual.dat and <prefix>.sav."""
        _check(lib().psst_case_solve(self._h))


class Solver:
    """A device context (``psst_ctx``) for a case: psstGridFV::solver_parameters +
    explicit_solve on the GPU."""

    def __init__(self, case, settings=None, **overrides):
        self.case = case
        s = settings if settings is not None else case.settings()
        for k, v in overrides.items():
            if k == "rk_coeff":
                for j, a in enumerate(v):
                    s.rk_coeff[j] = a
            else:
                setattr(s, k, v)
        self.settings = s
        self.dim = case.dim
        self.n_cells, self.n_ghosts = case.n_cells, case.n_ghosts
        m = case.mesh()
        h = C.c_void_p()
        st = lib().psst_ctx_create(C.byref(m), C.byref(s), C.byref(h))
        if st != 0:
            raise PsstError(st, lib().psst_last_error(None).decode())
        self._h = h
        _track(self)
        self.set_state(case.initial_state())

    def close(self):
        if getattr(self, "_h", None):
            lib().psst_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def set_state(self, prim):
        prim = np.ascontiguousarray(prim, dtype=np.float64)
        assert prim.shape == (self.dim + 3, self.n_cells + self.n_ghosts)
        _check(lib().psst_set_state(self._h, _dp(prim)), self._h)

    def state(self):
        prim = np.zeros((self.dim + 3, self.n_cells + self.n_ghosts))
        _check(lib().psst_get_state(self._h, _dp(prim)), self._h)
        return prim

    def conservative(self):
        w = np.zeros((self.dim + 2, self.n_cells))
        _check(lib().psst_get_conservative(self._h, _dp(w)), self._h)
        return w

    def run(self, n_iter, tol=0.0, allow_no_convergence=False):
        """Runs n_iter iterations; returns the residual history [n_iter, dim+2]
        (rows beyond the iterations actually run are zero when tol stops early)."""
        res = np.zeros((max(n_iter, 1), self.dim + 2))
        done = C.c_int32(0)
        st = lib().psst_run(self._h, int(n_iter), float(tol), _dp(res), C.byref(done))
        self.iters_done = int(done.value)
        if st == 4 and allow_no_convergence:
            return res[:n_iter]
        _check(st, self._h)
        return res[:n_iter]

    def set_residual_reference(self, res0):
        r = np.ascontiguousarray(res0, dtype=np.float64)
        _check(lib().psst_set_residual_reference(self._h, _dp(r)), self._h)

    def gradients(self):
        g = np.zeros((self.dim + 2, self.dim, self.n_cells))
        _check(lib().psst_debug_gradients(self._h, _dp(g)), self._h)
        return g

    def residual(self):
        r = np.zeros((self.dim + 2, self.n_cells))
        _check(lib().psst_debug_residual(self._h, _dp(r)), self._h)
        return r

    def kernel_timing(self, enable=True):
        _check(lib().psst_kernel_timing(self._h, int(bool(enable))), self._h)

    def kernel_stats(self, name):
        cnt = C.c_int64(0)
        ms = C.c_double(0)
        _check(lib().psst_kernel_stats(self._h, name.encode(), C.byref(cnt), C.byref(ms)), self._h)
        return int(cnt.value), float(ms.value)

    def set_precision(self, precision):
        """Switches the device state between FP64 and FP32 in place."""
        _check(lib().psst_set_precision(self._h, int(precision)), self._h)

    def kernel_names(self):
        buf = C.create_string_buffer(4096)
        _check(lib().psst_kernel_names(self._h, buf, 4096), self._h)
        return buf.value.decode().split(";")


class Batch:
    """Many independent cases advanced together on the GPU (``psst_batch_*``):
    one union mesh, one launch per kernel per stage for all instances, each
    instance frozen on the device once it reaches the tolerance.  Instances
    share the scheme (every ``psst_settings`` field except the boundary
    table) and may differ in mesh and boundary values.  In FP64 every
    instance is bitwise identical to its own ``Solver`` run."""

    def __init__(self, cases, settings=None, **overrides):
        self.cases = list(cases)
        n = len(self.cases)
        if n == 0:
            raise ValueError("empty batch")
        self.dim = self.cases[0].dim
        self._meshes = (MeshDesc * n)()
        self._settings = (Settings * n)()
        for i, c in enumerate(self.cases):
            self._meshes[i] = c.mesh()
            s = c.settings()
            if settings is not None:
                for name, _ in Settings._fields_:
                    if name not in ("bc", "n_bc"):
                        setattr(s, name, getattr(settings, name))
            for k, v in overrides.items():
                if k == "rk_coeff":
                    for j, a in enumerate(v):
                        s.rk_coeff[j] = a
                else:
                    setattr(s, k, v)
            self._settings[i] = s
        h = C.c_void_p()
        st = lib().psst_batch_create(n, self._meshes, self._settings, C.byref(h))
        if st != 0:
            raise PsstError(st, lib().psst_batch_last_error(None).decode())
        self._h = h
        _track(self)
        self.n = n
        self.precision = int(self._settings[0].precision)
        for i, c in enumerate(self.cases):
            self.set_state(i, c.initial_state())

    def _check(self, st):
        if st != 0:
            raise PsstError(st, lib().psst_batch_last_error(self._h).decode())

    def close(self):
        if getattr(self, "_h", None):
            lib().psst_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def set_state(self, i, prim):
        c = self.cases[i]
        prim = np.ascontiguousarray(prim, dtype=np.float64)
        assert prim.shape == (self.dim + 3, c.n_cells + c.n_ghosts)
        self._check(lib().psst_batch_set_state(self._h, int(i), _dp(prim)))

    def state(self, i):
        c = self.cases[i]
        prim = np.zeros((self.dim + 3, c.n_cells + c.n_ghosts))
        self._check(lib().psst_batch_get_state(self._h, int(i), _dp(prim)))
        return prim

    def set_residual_reference(self, i, res0):
        r = np.ascontiguousarray(res0, dtype=np.float64)
        self._check(lib().psst_batch_set_residual_reference(self._h, int(i), _dp(r)))

    def set_precision(self, precision):
        self._check(lib().psst_batch_set_precision(self._h, int(precision)))
        self.precision = int(precision)

    def set_enabled(self, i, enabled):
        """Excludes or includes instance i in later runs."""
        self._check(lib().psst_batch_set_enabled(self._h, int(i), int(bool(enabled))))

    def set_graph(self, enable):
        self._check(lib().psst_batch_set_graph(self._h, int(bool(enable))))

    def info(self):
        a = (C.c_int64 * 8)()
        self._check(lib().psst_batch_info(self._h, a))
        keys = ("instances", "n_cells", "n_ghosts", "n_faces", "tiles", "tile_cells", "tile_threads", "iterations")
        return dict(zip(keys, (int(x) for x in a)))

    def run(self, max_iter, tol=0.0, history=False, allow_no_convergence=False, allow_failures=False):
        """Runs every instance for up to max_iter iterations (each stops at its
        first iteration with max_j res_j/res0_j <= tol when tol > 0).  Returns a
        dict of per-instance ``iters``, ``status``, ``drop`` and, with
        history=True, ``hist`` [max_iter, n, dim+2].  Raises on any instance
        that did not converge (unless allow_no_convergence) or went
        non-finite (unless allow_failures); the statuses say which."""
        n = self.n
        iters = np.zeros(n, dtype=np.int32)
        status = np.zeros(n, dtype=np.int32)
        drop = np.zeros(n)
        hist = np.zeros((max(int(max_iter), 1), n, self.dim + 2)) if history else None
        st = lib().psst_batch_run(self._h, int(max_iter), float(tol), iters.ctypes.data_as(C.POINTER(C.c_int32)),
                                  status.ctypes.data_as(C.POINTER(C.c_int32)), _dp(drop),
                                  _dp(hist) if history else None)
        out = {"iters": iters, "status": status, "drop": drop}
        if history:
            out["hist"] = hist[:max_iter]
        if (st == 4 and allow_no_convergence) or (st in (3, 4) and allow_failures):
            return out
        self._check(st)
        return out

    def kernel_timing(self, enable=True):
        ctx = lib().psst_batch_context(self._h)
        _check(lib().psst_kernel_timing(ctx, int(bool(enable))), ctx)

    def kernel_stats(self, name):
        ctx = lib().psst_batch_context(self._h)
        cnt = C.c_int64(0)
        ms = C.c_double(0)
        _check(lib().psst_kernel_stats(ctx, name.encode(), C.byref(cnt), C.byref(ms)), ctx)
        return int(cnt.value), float(ms.value)


def autotune(case, settings=None, prim=None, trial_iters=10, **overrides):
    """Times the fused kernels' tile shapes on a case (``psst_autotune``) and
    returns (best (cells, threads), [((cells, threads), ms per iteration), ...])."""
    s = settings if settings is not None else case.settings()
    for k, v in overrides.items():
        setattr(s, k, v)
    m = case.mesh()
    prim = np.ascontiguousarray(case.initial_state() if prim is None else prim, dtype=np.float64)
    cand = (C.c_int32 * 64)()
    ms = (C.c_double * 32)()
    ncand = C.c_int32(0)
    best = (C.c_int32 * 2)()
    _check(lib().psst_autotune(C.byref(m), C.byref(s), _dp(prim), int(trial_iters), 32, cand, ms, C.byref(ncand),
                               best))
    table = [((cand[2 * k], cand[2 * k + 1]), ms[k]) for k in range(ncand.value)]
    return (best[0], best[1]), table


def sweep_schedule(case, chunk_cells=0, partition=0):
    """Builds and checks (on the host) the chunk schedule of the reference-ordered
    time step that a device context uses on a large mesh
    (``psst_debug_sweep_schedule``; partition 1 = Morton ranges, 2 = potential
    boxes, 0 = the context's choice); returns its figures as a dict."""
    m = case.mesh()
    st = (C.c_int64 * 10)()
    code = lib().psst_debug_sweep_schedule(C.byref(m), int(chunk_cells), int(partition), st)
    if code != 0:
        raise PsstError(code, lib().psst_last_error(None).decode())
    keys = ("levels", "chunks", "entries", "waits", "publishes", "widest_entry", "makespan_free",
            "makespan_wait4", "chunk_cells", "partition")
    return dict(zip(keys, (int(x) for x in st)))
