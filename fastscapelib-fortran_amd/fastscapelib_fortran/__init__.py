"""Drop-in for the f2py module ``fastscapelib_fortran`` (src_python/fastscapelib_fortran/__init__.py:1),
running FastScape_Execute_Step on an MI355X through the C ABI of libfastscape_hip.so.

The free functions keep the reference's call surface and f2py argument conventions
(src/FastScape_api.f90, usage as examples/FastScape_test.ipynb):

* scalars by value: ``fastscape_set_nx_ny(nx, ny)``, ``fastscape_set_dt(dt)``, ``fastscape_set_bc(ibc)``;
* ``intent(in)`` arrays are copied in: ``fastscape_init_h(h)``, ``fastscape_set_u(u)``, ...;
* copy-outs fill a caller-owned float64 array in place and return None: ``fastscape_copy_h(h)``;
* ``fastscape_get_step()`` and ``fastscape_get_sizes()`` return their values.

Documented divergences: call-order and size errors raise ``RuntimeError``/``ValueError`` where the
reference STOPs the process; ``fastscape_setup`` starts from h = 0 rather than Fortran random
numbers (every run sets h with ``fastscape_init_h``); routines outside the per-step update
(flexure) raise ``NotImplementedError``; ``fastscape_set_marine_parameters`` turns on the marine
transport (Marine) after Diffusion in every step, on the GPU; ``fastscape_strati`` tracks stratigraphic
reflectors on the GPU in every step and writes the reference's horizon files from the host at its output
steps (fastscapelib_fortran.strati); ``fastscape_vtk`` writes the
reference's legacy-VTK files on the host from the fields copied off the GPU; ``fastscape_set_v`` turns on the
advection of h, b and etot (Advect) at the start of every step, on the GPU.  The slope,
curvature, chi and catchment read-outs and the three fluxes are computed on the GPU from the resident
state; the catchment labels a catchment (outlet + 1) / nn where the reference draws a random number.

``Batch`` runs many independent landscapes of different sizes in one set of kernel launches.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import FastscapeError, check, host_f64, indexx, lib, ptr

__all__ = [
    "fastscape_init", "fastscape_set_nx_ny", "fastscape_setup", "fastscape_set_xl_yl", "fastscape_set_dt",
    "fastscape_init_h", "fastscape_set_erosional_parameters", "fastscape_set_u", "fastscape_set_bc",
    "fastscape_execute_step", "fastscape_copy_h", "fastscape_copy_basement", "fastscape_copy_total_erosion",
    "fastscape_copy_erosion_rate", "fastscape_copy_drainage_area", "fastscape_copy_lake_depth",
    "fastscape_copy_slope", "fastscape_copy_curvature", "fastscape_copy_chi", "fastscape_copy_catchment",
    "fastscape_get_fluxes", "fastscape_set_v", "fastscape_vtk", "fastscape_set_marine_parameters",
    "fastscape_init_f", "fastscape_copy_f", "fastscape_strati",
    "fastscape_get_step", "fastscape_get_sizes", "fastscape_destroy", "fastscape_set_h", "fastscape_set_basement",
    "fastscape_set_precip", "fastscape_set_all_layers", "fastscape_reset_cumulative_erosion", "fastscape_view",
    "fastscape_debug", "Batch", "Lanes", "Pool", "FastscapeError", "indexx",
]


class Batch:
    """Independent landscapes resident in HBM, stepped together.

    ``sizes`` is a sequence of (nx, ny).  Setters take the instance id first and otherwise mirror
    the fastscape_* functions.  Arrays may be numpy (host) or torch CUDA tensors (device, no copy
    through the host).
    """

    def __init__(self, sizes):
        sizes = [(int(a), int(b)) for a, b in sizes]
        if not sizes:
            raise ValueError("Batch needs at least one landscape")
        n = len(sizes)
        nx = (C.c_int * n)(*[s[0] for s in sizes])
        ny = (C.c_int * n)(*[s[1] for s in sizes])
        h = C.c_void_p()
        check(lib.fsx_create(n, nx, ny, C.byref(h)), "fsx_create")
        self._h = h
        self.sizes = dict(enumerate(sizes))
        self.nn = {k: a * b for k, (a, b) in self.sizes.items()}
        self.slots = n

    # -- lifetime --
    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            lib.fsx_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __len__(self):
        return len(self.sizes)

    # -- helpers --
    def _src(self, k, x, name):
        """(pointer, on_device, keepalive) for an input array of instance k."""
        if hasattr(x, "data_ptr") and getattr(x, "is_cuda", False):
            import torch
            t = x.reshape(-1)
            if t.dtype != torch.float64:
                t = t.to(torch.float64)
            t = t.contiguous()
            if t.numel() < self.nn[k]:
                raise ValueError(f"{name}: expected at least {self.nn[k]} values")
            return C.c_void_p(t.data_ptr()), 1, t
        a = host_f64(x, self.nn[k], name)
        return ptr(a), 0, a

    # -- setters --
    def set_xl_yl(self, k, xl, yl):
        check(lib.fsx_set_xl_yl(self._h, k, float(xl), float(yl)), "set_xl_yl")
        self.__dict__.setdefault("_xlyl", {})[k] = (float(xl), float(yl))  # Pool has its own __init__
    def set_dt(self, k, dt):
        check(lib.fsx_set_dt(self._h, k, float(dt)), "set_dt")
        self.__dict__.setdefault("_dt", {})[k] = float(dt)
    def set_bc(self, k, ibc): check(lib.fsx_set_bc(self._h, k, int(ibc)), "set_bc")
    def set_newton_tol(self, k, tol): check(lib.fsx_set_newton_tol(self._h, k, float(tol)), "set_newton_tol")

    def set_marine_parameters(self, k, sealevel, poro1, poro2, zporo1, zporo2, ratio, layer, kdsea1, kdsea2):
        """Marine transport of landscape k after Diffusion in every step (FastScape_Set_Marine_Parameters)."""
        check(lib.fsx_set_marine_parameters(self._h, k, *[float(x) for x in (sealevel, poro1, poro2, zporo1, zporo2,
                                                                          ratio, layer, kdsea1, kdsea2)]),
              "set_marine_parameters")
        self.__dict__.setdefault("_sealevel", {})[k] = float(sealevel)

    def strati(self, k, nstep, nreflector, nfreq, vex, directory="VTK"):
        """Stratigraphic horizons of landscape k (FastScape_Strati): nreflector reflectors, a new one active
        every nstep // nreflector steps, the reference's VTK files written into ``directory`` every nfreq
        steps (fastscapelib_fortran.strati)."""
        check(lib.fsx_strati(self._h, k, int(nstep), int(nreflector), int(nfreq), float(vex)), "strati")
        nref = int(nreflector)
        self.__dict__.setdefault("_strati", {})[k] = dict(
            nref=nref, nfreq=int(nfreq), nfreqref=int(nstep) // nref, iref=1, vex=float(vex), directory=directory,
            last=None, fields=np.zeros((nref, 10, self.nn[k])))

    def strati_read(self, k, which, rows):
        out = np.empty((rows, self.nn[k]))
        check(lib.fsx_strati_read(self._h, k, int(which), out.ctypes.data_as(C.POINTER(C.c_double))), "strati_read")
        return out

    def _strati_output(self, k, step):
        """One output step of Strati (src/FastScape_ctx.f90:832-836; src/Strati.f90:1-82) for landscape k."""
        from . import strati as S
        st = self._strati[k]
        if step % st["nfreqref"] == 0:
            st["iref"] += 1
            check(lib.fsx_strati_set_active(self._h, k, st["iref"]), "strati_set_active")
        nref = st["nref"]
        refl = self.strati_read(k, _lib.FSX_STRATI_REFLECTORS, nref)
        st["fields"][:, 9] = self.strati_read(k, _lib.FSX_STRATI_ERODED, nref)
        slopes = self.strati_read(k, _lib.FSX_STRATI_SLOPES, nref)
        dist = self.strati_read(k, _lib.FSX_STRATI_SHORE, nref + 1)
        b = self.field(k, _lib.FSX_B)
        dt = self.__dict__.get("_dt", {}).get(k, 0.0)
        S.update_fields(st["fields"], refl, slopes, dist[:nref], b, self.field(k, _lib.FSX_FMIX), st["iref"],
                        self.__dict__.get("_sealevel", {}).get(k, 0.0), dt * st["nfreqref"])
        nx, ny = self.sizes[k]
        xl, yl = self.__dict__.get("_xlyl", {}).get(k, (float(nx - 1), float(ny - 1)))
        S.write_output(step, nx, ny, xl, yl, refl, st["fields"], b, dist[nref], st["vex"], st["iref"],
                       st["directory"])
        st["last"] = step
    def set_active(self, k, on): check(lib.fsx_set_active(self._h, k, 1 if on else 0), "set_active")

    def set_erosional_parameters(self, k, kf, kfsed, m, n, kd, kdsed, g1, g2, p):
        pkf, dkf, a1 = self._src(k, kf, "kf")
        pkd, dkd, a2 = self._src(k, kd, "kd")
        if dkf != dkd:
            a1 = host_f64(kf.cpu().numpy() if dkf else kf, self.nn[k], "kf")
            a2 = host_f64(kd.cpu().numpy() if dkd else kd, self.nn[k], "kd")
            pkf, pkd, dkf = ptr(a1), ptr(a2), 0
        check(lib.fsx_set_erosional_parameters(self._h, k, pkf, float(kfsed), float(m), float(n), pkd, float(kdsed),
                                               float(g1), float(g2), float(p), dkf), "set_erosional_parameters")

    def init_h(self, k, h):
        p, d, _a = self._src(k, h, "h")
        check(lib.fsx_init_h(self._h, k, p, d), "init_h")

    def set_v(self, k, vx, vy):
        """Velocity field of landscape k (FastScape_Set_V): advection of h, b and etot starts each step."""
        p1, d1, a1 = self._src(k, vx, "vx")
        p2, d2, a2 = self._src(k, vy, "vy")
        if d1 != d2:
            a1 = host_f64(vx.cpu().numpy() if d1 else vx, self.nn[k], "vx")
            a2 = host_f64(vy.cpu().numpy() if d2 else vy, self.nn[k], "vy")
            p1, p2, d1 = ptr(a1), ptr(a2), 0
        check(lib.fsx_set_v(self._h, k, p1, p2, d1), "set_v")

    def set_field(self, k, field, x, name="field"):
        p, d, _a = self._src(k, x, name)
        check(lib.fsx_set_field(self._h, k, field, p, d), "set_field")

    def set_u(self, k, u): self.set_field(k, _lib.FSX_U, u, "u")
    def set_h(self, k, h): self.set_field(k, _lib.FSX_H, h, "h")
    def set_basement(self, k, b): self.set_field(k, _lib.FSX_B, b, "b")
    def set_precip(self, k, p): self.set_field(k, _lib.FSX_PRECIP, p, "precip")
    def init_f(self, k, f): self.set_field(k, _lib.FSX_FMIX, f, "f")

    def set_all_layers(self, k, dh):
        p, d, _a = self._src(k, dh, "dh")
        check(lib.fsx_set_all_layers(self._h, k, p, d), "set_all_layers")

    def reset_cumulative_erosion(self, k): check(lib.fsx_reset_cumulative_erosion(self._h, k), "reset")

    # -- stepping --
    def step(self, nsteps=1, stream=None, sync=True):
        s = C.c_void_p(stream) if isinstance(stream, int) else (stream or C.c_void_p())
        st = self.__dict__.get("_strati")
        if not st:
            check(lib.fsx_execute_steps(self._h, int(nsteps), s, 1 if sync else 0), "execute_steps")
            return
        # with stratigraphy on, the steps run in segments that end at output steps, where the files are
        # written (the reference writes them from inside its step, src/FastScape_ctx.f90:832-836)
        left = int(nsteps)
        while left > 0:
            n = min([left] + [q["nfreq"] - self.get_step(k) % q["nfreq"] for k, q in st.items()])
            check(lib.fsx_execute_steps(self._h, n, s, 1), "execute_steps")
            left -= n
            for k, q in st.items():
                t = self.get_step(k)
                if t % q["nfreq"] == 0 and t != q["last"]:
                    self._strati_output(k, t)
        if nsteps == 0:
            check(lib.fsx_execute_steps(self._h, 0, s, 1 if sync else 0), "execute_steps")

    def synchronize(self): check(lib.fsx_synchronize(self._h), "synchronize")

    def time_phase(self, name):
        """Time one phase (a name of ``_lib.PHASES``, or None) live while steps replay HIP graphs."""
        k = -1 if name is None else _lib.PHASES.index(name)
        check(lib.fsx_time_phase(self._h, k), "time_phase")

    TUNING = ("walk_gx", "jump_blocks", "tour_gx", "nr_small_gx", "area_wg", "bor_gx", "flow_wg", "flow_split",
              "walk_big", "mfd_split", "tile_h", "flow_nk")

    def autotune(self, reps=2):
        """Choose the launch shapes of the routing kernels and of the dataflow stream-power sweep by
        timing candidates on the current landscapes (whose elevations it leaves unchanged); returns the
        chosen shapes."""
        out = (C.c_int * len(self.TUNING))()
        check(lib.fsx_autotune(self._h, int(reps), out), "autotune")
        return dict(zip(self.TUNING, list(out)))

    def tuning(self):
        out = (C.c_int * len(self.TUNING))()
        check(lib.fsx_get_tuning(self._h, out), "get_tuning")
        return dict(zip(self.TUNING, list(out)))

    def set_tuning(self, shapes):
        """Set the launch shapes (a dict as returned by ``autotune``/``tuning``; missing keys keep
        their current value)."""
        cur = self.tuning()
        cur.update({k: int(v) for k, v in dict(shapes).items() if k in self.TUNING})
        arr = (C.c_int * len(self.TUNING))(*[cur[k] for k in self.TUNING])
        check(lib.fsx_set_tuning(self._h, arr), "set_tuning")

    def graphs(self, on=True):
        """Replay each step as one captured HIP graph (default) or launch its kernels one by one."""
        check(lib.fsx_set_graphs(self._h, 1 if on else 0), "set_graphs")

    def route(self, sync=True):
        check(lib.fsx_route(self._h, C.c_void_p(), 1 if sync else 0), "route")

    # -- read-outs --
    def field(self, k, field, out=None):
        if out is None:
            out = np.empty(self.nn[k], np.float64)
        if hasattr(out, "data_ptr") and getattr(out, "is_cuda", False):
            check(lib.fsx_copy_field(self._h, k, field, C.c_void_p(out.data_ptr()), 1), "copy_field")
            return out
        if out.dtype != np.float64 or not out.flags.c_contiguous or out.size < self.nn[k]:
            raise ValueError(f"copy target must be a contiguous float64 array of >= {self.nn[k]} values")
        check(lib.fsx_copy_field(self._h, k, field, ptr(out), 0), "copy_field")
        return out

    def derived(self, k, which, out=None):
        """Slope (FSX_SLOPE), curvature (FSX_CURVATURE) or chi (FSX_CHI) of landscape k, computed on the
        GPU from its current state (FastScape_Copy_Slope / _Curvature / _Chi)."""
        if out is None:
            out = np.empty(self.nn[k], np.float64)
        if hasattr(out, "data_ptr") and getattr(out, "is_cuda", False):
            check(lib.fsx_copy_derived(self._h, k, which, C.c_void_p(out.data_ptr()), 1), "copy_derived")
            return out
        if out.dtype != np.float64 or not out.flags.c_contiguous or out.size < self.nn[k]:
            raise ValueError(f"copy target must be a contiguous float64 array of >= {self.nn[k]} values")
        check(lib.fsx_copy_derived(self._h, k, which, ptr(out), 0), "copy_derived")
        return out

    def slope(self, k): return self.derived(k, _lib.FSX_SLOPE)
    def curvature(self, k): return self.derived(k, _lib.FSX_CURVATURE)
    def chi(self, k): return self.derived(k, _lib.FSX_CHI)
    def catchment(self, k): return self.derived(k, _lib.FSX_CATCHMENT)

    def derived_all(self, which, ks=None):
        """One read-out (FSX_SLOPE, FSX_CURVATURE, FSX_CHI or FSX_CATCHMENT) of the landscapes ks (every
        landscape held, by id, by default) in one launch: a list of arrays in the order of ks."""
        ks = sorted(self.nn) if ks is None else [int(k) for k in ks]
        for k in ks:
            if k not in self.nn:
                raise ValueError(f"derived_all: no landscape {k}")
        outs = [np.empty(self.nn[k], np.float64) for k in ks]
        ptrs = (C.c_void_p * self.slots)()  # the library reads one entry per slot
        for k, o in zip(ks, outs):
            ptrs[k] = o.ctypes.data
        check(lib.fsx_copy_derived_batch(self._h, which, ptrs, 0), "copy_derived_batch")
        return outs

    def fluxes_all(self, ks=None):
        """(tectonic, erosion, boundary) flux of the landscapes ks (every landscape held, by id, by
        default) in one launch: an array [len(ks), 3] in the order of ks."""
        ks = sorted(self.nn) if ks is None else [int(k) for k in ks]
        for k in ks:
            if k not in self.nn:
                raise ValueError(f"fluxes_all: no landscape {k}")
        out = np.empty((self.slots, 3), np.float64)  # the library writes three values per slot
        check(lib.fsx_get_fluxes_batch(self._h, ptr(out)), "get_fluxes_batch")
        return out[ks] if ks else np.empty((0, 3), np.float64)

    def vtk(self, k, f, vex, directory="VTK"):
        """FastScape_VTK for landscape k: legacy-VTK files of h (and b, sea level for vex < 0) with the
        caller's field f, written on the host from the fields copied off the GPU."""
        from .vtk import write_vtk
        nx, ny = self.sizes[k]
        xl, yl = self.__dict__.get("_xlyl", {}).get(k, (float(nx - 1), float(ny - 1)))  # library default dx = 1
        return write_vtk(self.get_step(k), nx, ny, xl, yl, self.h(k), self.field(k, _lib.FSX_B), f, float(vex),
                         0.0, directory)

    def fluxes(self, k):
        """(tectonic, erosion, boundary) flux of landscape k (FastScape_Get_Fluxes)."""
        o = (C.c_double * 3)()
        check(lib.fsx_get_fluxes(self._h, k, o), "get_fluxes")
        return o[0], o[1], o[2]

    def ifield(self, k, field):
        n = self.nn[k] * (8 if field in (_lib.FSX_DON, _lib.FSX_MREC) else 1)
        out = np.empty(n, np.int32)
        check(lib.fsx_copy_ifield(self._h, k, field, ptr(out), 0), "copy_ifield")
        return out.reshape(self.nn[k], 8) if n != self.nn[k] else out

    def h(self, k): return self.field(k, _lib.FSX_H)
    def rec(self, k): return self.ifield(k, _lib.FSX_REC)
    def donors(self, k): return self.ifield(k, _lib.FSX_NDON), self.ifield(k, _lib.FSX_DON)
    def order(self, k): return self.ifield(k, _lib.FSX_ORDER)

    def mrec(self, k):
        """Multiple-flow receivers of landscape k: (mnrec[nn], mrec[nn, 8]) in the reference's slot order."""
        return self.ifield(k, _lib.FSX_MNREC), self.ifield(k, _lib.FSX_MREC)

    def morder(self, k):
        """Multiple-flow order of landscape k: every node before its receivers (the role of mstack)."""
        return self.ifield(k, _lib.FSX_MSTACK)

    def get_step(self, k):
        s = C.c_int()
        check(lib.fsx_get_step(self._h, k, C.byref(s)), "get_step")
        return s.value

    def stats(self, k):
        o = (C.c_longlong * 6)()
        check(lib.fsx_get_stats(self._h, k, o), "get_stats")
        return dict(zip(["nbasins", "nconn", "ntree", "levels", "ngs", "newton_iters"], list(o)))

    def profiling(self, on=True):
        """Record a HIP event pair around every phase of each step (on the stream the step runs on)."""
        check(lib.fsx_set_profiling(self._h, 1 if on else 0), "set_profiling")

    def reset_timers(self): check(lib.fsx_reset_timers(self._h), "reset_timers")

    def timers(self):
        """{phase: (total ms, launches)} accumulated while profiling was on."""
        n = len(_lib.PHASES)
        ms = (C.c_double * n)()
        cnt = (C.c_longlong * n)()
        rc = lib.fsx_get_timers(self._h, ms, cnt, n)
        if rc < 0:
            check(rc, "get_timers")
        return {p: (ms[i], cnt[i]) for i, p in enumerate(_lib.PHASES)}

    def memory(self):
        a, b = C.c_longlong(), C.c_longlong()
        check(lib.fsx_memory(self._h, C.byref(a), C.byref(b)), "memory")
        return a.value, b.value


class Lanes:
    """Landscapes split over ``lanes`` Batch contexts, each stepping on its own HIP stream.

    Per-landscape kernels (level sweeps, edge sort, basin walk, spanning forest) run one workgroup
    per landscape, so a batch of a few dozen landscapes leaves most compute units idle while they
    run.  ``step`` enqueues every lane's steps before waiting for any, so one lane's per-landscape
    kernels overlap the other lanes' node-parallel ones.  Landscapes are dealt to lanes by size
    (largest first, to the lane with the fewest cells); results do not depend on the split.  The
    interface is Batch's: per-landscape calls take the landscape id of the whole set.
    """

    _PER = {"set_xl_yl", "set_dt", "set_bc", "set_newton_tol", "set_active", "set_erosional_parameters",
            "init_h", "set_field", "set_u", "set_v", "set_h", "set_basement", "set_precip", "set_all_layers",
            "reset_cumulative_erosion", "field", "ifield", "h", "rec", "donors", "order", "mrec", "morder",
            "get_step", "stats", "vtk", "derived", "slope", "curvature", "chi", "catchment", "fluxes",
            "set_marine_parameters", "init_f", "strati"}
    TUNING = Batch.TUNING

    @staticmethod
    def deal(sizes, lanes):
        """Landscape ids per lane: largest landscape first, each to the lane with the fewest cells so
        far (ties: the lower lane); ids ascending within a lane.  At most one lane per landscape."""
        sizes = [(int(a), int(b)) for a, b in sizes]
        nl = max(1, min(int(lanes), len(sizes)))
        load, members = [0] * nl, [[] for _ in range(nl)]
        for k in sorted(range(len(sizes)), key=lambda k: (-sizes[k][0] * sizes[k][1], k)):
            j = min(range(nl), key=lambda j: (load[j], j))
            members[j].append(k)
            load[j] += sizes[k][0] * sizes[k][1]
        return [sorted(m) for m in members]

    def __init__(self, sizes, lanes=2):
        sizes = [(int(a), int(b)) for a, b in sizes]
        if not sizes:
            raise ValueError("Lanes needs at least one landscape")
        self.sizes = dict(enumerate(sizes))
        self.nn = {k: a * b for k, (a, b) in self.sizes.items()}
        self.members = Lanes.deal(sizes, lanes)
        self.where = {k: (j, i) for j, m in enumerate(self.members) for i, k in enumerate(m)}
        self.lanes = []
        try:
            for m in self.members:
                self.lanes.append(Batch([sizes[k] for k in m]))
        except Exception:
            self.close()
            raise

    def __getattr__(self, name):
        if name in Lanes._PER:
            def per(k, *a, **kw):
                j, i = self.where[k]
                return getattr(self.lanes[j], name)(i, *a, **kw)
            return per
        raise AttributeError(name)

    def derived_all(self, which, ks=None):
        """Batch.derived_all over the lanes: one launch per lane, results in the order of ks."""
        ks = list(range(len(self.sizes))) if ks is None else list(ks)
        got = {}
        for j, b in enumerate(self.lanes):
            mine = [k for k in ks if self.where[k][0] == j]
            if mine:
                for k, o in zip(mine, b.derived_all(which, [self.where[k][1] for k in mine])):
                    got[k] = o
        return [got[k] for k in ks]

    def fluxes_all(self):
        """Batch.fluxes_all over the lanes: an array [n, 3] in landscape order."""
        out = np.empty((len(self.sizes), 3), np.float64)
        for j, b in enumerate(self.lanes):
            fl = b.fluxes_all()
            for i, k in enumerate(self.members[j]):
                out[k] = fl[i]
        return out

    def __len__(self):
        return len(self.sizes)

    def close(self):
        for b in getattr(self, "lanes", []):
            b.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def step(self, nsteps=1, sync=True):
        # one step per lane in turn: a lane whose host side waits (its phase-timer ring filling up)
        # still has the other lanes' steps queued on the device
        for _ in range(int(nsteps)):
            for b in self.lanes:
                b.step(1, sync=False)
        if sync:
            for b in self.lanes:
                b.step(0, sync=True)  # waits for the lane's stream and reports its errors

    def synchronize(self):
        for b in self.lanes:
            b.step(0, sync=True)

    def _all(self, name, *a):
        return [getattr(b, name)(*a) for b in self.lanes]

    def graphs(self, on=True): self._all("graphs", on)
    def profiling(self, on=True): self._all("profiling", on)
    def reset_timers(self): self._all("reset_timers")
    def time_phase(self, name): self._all("time_phase", name)
    def set_tuning(self, shapes): self._all("set_tuning", shapes)
    def route(self, sync=True): self._all("route", sync)

    def autotune(self, reps=2):
        """Tunes each lane on its own landscapes (lanes run one after another while tuning)."""
        return self._all("autotune", reps)

    def tuning(self):
        return self._all("tuning")

    def timers(self, lane=None):
        """{phase: (ms, launches)}: of one lane, or summed over lanes (lanes run concurrently, so the
        sum over-counts wall time)."""
        if lane is not None:
            return self.lanes[lane].timers()
        out = {}
        for b in self.lanes:
            for p, (ms, c) in b.timers().items():
                m0, c0 = out.get(p, (0.0, 0))
                out[p] = (m0 + ms, c0 + c)
        return out

    def memory(self):
        tot = [b.memory() for b in self.lanes]
        return sum(t[0] for t in tot), tot[0][1]


class Pool(Batch):
    """A Batch whose landscapes come and go: an arena of ``capacity_nodes`` nodes and ``slots``
    landscape slots in HBM.  ``add(nx, ny)`` places a landscape (SetUp defaults) and returns its id,
    or None when no slot or contiguous node range is free; ``remove(id)`` frees it.  ``step`` advances
    every landscape currently held."""

    def __init__(self, capacity_nodes, slots=256):
        h = C.c_void_p()
        check(lib.fsx_create_pool(int(capacity_nodes), int(slots), C.byref(h)), "fsx_create_pool")
        self._h = h
        self.sizes = {}
        self.nn = {}
        self.capacity = int(capacity_nodes)
        self.slots = int(slots)

    def add(self, nx, ny):
        k = C.c_int(-1)
        rc = lib.fsx_add_instance(self._h, int(nx), int(ny), C.byref(k))
        if rc == _lib.FSX_ENOMEM:
            return None
        check(rc, "add_instance")
        self.sizes[k.value] = (int(nx), int(ny))
        self.nn[k.value] = int(nx) * int(ny)
        return k.value

    def remove(self, k):
        check(lib.fsx_remove_instance(self._h, int(k)), "remove_instance")
        del self.sizes[k]
        del self.nn[k]
        for d in ("_xlyl", "_dt", "_sealevel", "_strati"):
            self.__dict__.get(d, {}).pop(k, None)

    def ready(self):
        """True once every step queued on the pool's stream has finished (errors raised as after a
        synchronous step), False while some still run."""
        rc = lib.fsx_query(self._h)
        if rc < 0:
            check(rc, "query")
        return rc == 1

    def info(self):
        """(free nodes, largest free contiguous range, free slots)."""
        a, b, c = C.c_longlong(), C.c_longlong(), C.c_int()
        check(lib.fsx_pool_info(self._h, C.byref(a), C.byref(b), C.byref(c)), "pool_info")
        return a.value, b.value, c.value


# ---- the f2py-style module state (one landscape, as the reference's FastScapeContext) ----------
class _State:
    def __init__(self):
        self.nx = 0
        self.ny = 0
        self.batch: Batch | None = None


_S = _State()


def _need():
    if _S.batch is None:
        raise RuntimeError("You need to run SetUp first")
    return _S.batch


def fastscape_init():
    """FastScape_Init (src/FastScape_api.f90:146-156)."""
    if _S.batch is not None:
        _S.batch.close()
    _S.__init__()


def fastscape_set_nx_ny(nx, ny):
    """FastScape_Set_NX_NY (src/FastScape_api.f90:475-487)."""
    _S.nx, _S.ny = int(nx), int(ny)


def fastscape_setup():
    """FastScape_Setup (src/FastScape_api.f90:160-170; SetUp src/FastScape_ctx.f90:57-113)."""
    if _S.nx == 0:
        raise RuntimeError("FastScapeSetup - You need to set nx first")
    if _S.ny == 0:
        raise RuntimeError("FastScapeSetup - You need to set ny first")
    if _S.batch is not None:
        _S.batch.close()
    _S.batch = Batch([(_S.nx, _S.ny)])


def fastscape_destroy():
    """FastScape_Destroy (src/FastScape_api.f90:174-184)."""
    if _S.batch is not None:
        _S.batch.close()
        _S.batch = None


def fastscape_set_xl_yl(xl, yl): _need().set_xl_yl(0, xl, yl)
def fastscape_set_dt(dt): _need().set_dt(0, dt)
def fastscape_set_bc(ibc): _need().set_bc(0, ibc)
def fastscape_init_h(h): _need().init_h(0, h)
def fastscape_set_u(u): _need().set_u(0, u)
def fastscape_set_v(ux, uy): _need().set_v(0, ux, uy)
def fastscape_set_h(h): _need().set_h(0, h)
def fastscape_set_basement(b): _need().set_basement(0, b)
def fastscape_set_precip(p): _need().set_precip(0, p)
def fastscape_set_all_layers(dh): _need().set_all_layers(0, dh)
def fastscape_init_f(f): _need().init_f(0, f)


def fastscape_strati(nstepp, nreflectorp, nfreqp, vexp):
    """FastScape_Strati (src/FastScape_api.f90:745-758): stratigraphic horizons written into VTK/ every
    nfreqp steps (Horizon*.vtk, Strati-*.vtk, CUBE.vtk), as the reference's Execute_Step does."""
    _need().strati(0, nstepp, nreflectorp, nfreqp, vexp)
def fastscape_reset_cumulative_erosion(): _need().reset_cumulative_erosion(0)


def fastscape_set_erosional_parameters(kf, kfsed, m, n, kd, kdsed, g1, g2, p):
    """FastScape_Set_Erosional_Parameters (src/FastScape_api.f90:523-536)."""
    _need().set_erosional_parameters(0, kf, kfsed, m, n, kd, kdsed, g1, g2, p)


def fastscape_set_marine_parameters(sl, p1, p2, z1, z2, r, l, kds1, kds2):
    """FastScape_Set_Marine_Parameters (src/FastScape_api.f90:540-552): sea level, surface porosities and
    e-folding depths of silt and sand, silt ratio, mixed-layer thickness, marine diffusivities."""
    _need().set_marine_parameters(0, sl, p1, p2, z1, z2, r, l, kds1, kds2)


def fastscape_execute_step():
    """FastScape_Execute_Step (src/FastScape_api.f90:201-263)."""
    _need().step(1)


def _copy_into(field, out):
    b = _need()
    if not isinstance(out, np.ndarray) or out.dtype != np.float64 or not out.flags.c_contiguous:
        raise ValueError("copy target must be a contiguous float64 numpy array (f2py intent(inout))")
    if out.size < b.nn[0]:
        raise ValueError(f"copy target holds {out.size} values, need {b.nn[0]}")
    b.field(0, field, out.reshape(-1))


def fastscape_copy_h(h): _copy_into(_lib.FSX_H, h)
def fastscape_copy_basement(b): _copy_into(_lib.FSX_B, b)
def fastscape_copy_total_erosion(e): _copy_into(_lib.FSX_ETOT, e)
def fastscape_copy_erosion_rate(e): _copy_into(_lib.FSX_ERATE, e)
def fastscape_copy_drainage_area(a): _copy_into(_lib.FSX_AREA, a)
def fastscape_copy_lake_depth(d): _copy_into(_lib.FSX_LAKE_DEPTH, d)
def fastscape_copy_f(f): _copy_into(_lib.FSX_FMIX, f)


def _derived_into(which, out):
    b = _need()
    if not isinstance(out, np.ndarray) or out.dtype != np.float64 or not out.flags.c_contiguous:
        raise ValueError("copy target must be a contiguous float64 numpy array (f2py intent(inout))")
    if out.size < b.nn[0]:
        raise ValueError(f"copy target holds {out.size} values, need {b.nn[0]}")
    b.derived(0, which, out.reshape(-1))


def fastscape_copy_slope(s):
    """FastScape_Copy_Slope (src/FastScape_api.f90:395-407): slope in degrees, outer ring 0."""
    _derived_into(_lib.FSX_SLOPE, s)


def fastscape_copy_curvature(c):
    """FastScape_Copy_Curvature (src/FastScape_api.f90:411-423), outer ring 0."""
    _derived_into(_lib.FSX_CURVATURE, c)


def fastscape_copy_chi(c):
    """FastScape_Copy_Chi (src/FastScape_api.f90:379-391), along the last step's receivers."""
    _derived_into(_lib.FSX_CHI, c)


def fastscape_vtk(fp, vexp):
    """FastScape_VTK (src/FastScape_api.f90:728-741): VTK/Topography<step>.vtk (and Basement, SeaLevel
    for vexp < 0) in the working directory, byte for byte as Make_VTK writes them."""
    _need().vtk(0, fp, vexp)


def fastscape_copy_catchment(c):
    """FastScape_Copy_Catchment (src/FastScape_api.f90:427-439): (outlet + 1) / nn of each node's tree
    before closed-basin rerouting; the reference gives that tree a random number instead."""
    _derived_into(_lib.FSX_CATCHMENT, c)


def fastscape_get_fluxes():
    """FastScape_Get_Fluxes (src/FastScape_api.f90:762-774): (tectonic, erosion, boundary) flux, as
    f2py returns the three intent(out) scalars."""
    return _need().fluxes(0)


def fastscape_get_step():
    """FastScape_Get_Step (src/FastScape_api.f90:572-584)."""
    return _need().get_step(0)


def fastscape_get_sizes():
    """FastScape_Get_Sizes (src/FastScape_api.f90:556-568)."""
    return _S.nx, _S.ny


def fastscape_view():
    """FastScape_View (src/FastScape_ctx.f90:378-393), summary of the resident state."""
    b = _need()
    h = b.h(0)
    print("FastScapeContext:")
    print("nx,ny", _S.nx, _S.ny)
    print("nn", _S.nx * _S.ny)
    print("step", b.get_step(0))
    print("h", h.min(), h.mean(), h.max())


def fastscape_debug():
    """FastScape_Debug (src/FastScape_ctx.f90:513-561): routing statistics of the last step."""
    b = _need()
    st = b.stats(0)
    rec = b.rec(0)
    print("--------------------------------------------------------")
    print("Time step", b.get_step(0))
    print("Total number of nodes (nx*ny)", _S.nx * _S.ny)
    print("Total number of self donors", int((rec == np.arange(rec.size)).sum()))
    print("Number of basins / basin-graph edges / MST edges", st["nbasins"], st["nconn"], st["ntree"])
    print("Number of Gauss-Siedel iterations (SPL)", st["ngs"])
