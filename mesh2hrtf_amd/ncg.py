"""ctypes binding of libncg (include/ncg.h) — the GPU BEM solve behind the drop-in NumCalc.

The library is the product path; there is no CPU fallback.  Loading fails loudly when
the built ``libncg.so`` is missing (build it with ``__graft_entry__.build()`` or
``make -C mesh2hrtf_amd/csrc``).
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libncg.so")
NUMCALC = os.path.join(PKG, "bin", "NumCalc")

FP64, MIXED = 0, 1
CGS, GMRES, LU = 0, 1, 4
ILU, SCALING, NONE, BLOCK = 0, 1, 2, 3
BC_VELOCITY, BC_PRESSURE, BC_ADMITTANCE = 0, 1, 2


class NcgError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("libncg error %d: %s" % (code, msg))
        self.code = code


class _Method(ctypes.Structure):
    _fields_ = [("method_bem", ctypes.c_int32), ("min_expansion_terms", ctypes.c_int32),
                ("cluster_edge", ctypes.c_double)]


class _Geometry(ctypes.Structure):
    _fields_ = [("n_elements", ctypes.c_int32), ("nodes_per_element", ctypes.c_void_p),
                ("node_xyz", ctypes.c_void_p), ("bc_kind", ctypes.c_void_p), ("n_eval", ctypes.c_int32),
                ("eval_xyz", ctypes.c_void_p), ("harmonic_time_factor", ctypes.c_double),
                ("interior", ctypes.c_int32), ("node_index", ctypes.c_void_p), ("element_group", ctypes.c_void_p),
                ("method", _Method)]


class _FmmInfo(ctypes.Structure):
    _fields_ = [("method", ctypes.c_int32), ("levels", ctypes.c_int32), ("clusters", ctypes.c_int32 * 10),
                ("radius_max", ctypes.c_double * 10), ("radius_min", ctypes.c_double * 10),
                ("expansion_terms", ctypes.c_int32 * 10), ("points_theta", ctypes.c_int32 * 10),
                ("points_phi", ctypes.c_int32 * 10), ("points_sphere", ctypes.c_int32 * 10),
                ("nnz_near", ctypes.c_int64), ("nnz_t", ctypes.c_int64 * 10), ("nnz_d", ctypes.c_int64 * 10),
                ("nnz_s", ctypes.c_int64 * 10), ("canceled_pairs", ctypes.c_int64 * 10),
                ("far_pairs", ctypes.c_int64 * 10), ("eval_clusters", ctypes.c_int32),
                ("eval_radius_max", ctypes.c_double), ("eval_radius_min", ctypes.c_double),
                ("warning", ctypes.c_char * 256)]


def _info_dict(i):
    L = i.levels
    return dict(method=i.method, levels=L, clusters=list(i.clusters[:L]), radius_max=list(i.radius_max[:L]),
                radius_min=list(i.radius_min[:L]), expansion_terms=list(i.expansion_terms[:L]),
                points_theta=list(i.points_theta[:L]), points_phi=list(i.points_phi[:L]),
                points_sphere=list(i.points_sphere[:L]), nnz_near=i.nnz_near, nnz_t=list(i.nnz_t[:L]),
                nnz_d=list(i.nnz_d[:L]), nnz_s=list(i.nnz_s[:L]), canceled_pairs=list(i.canceled_pairs[:L]),
                far_pairs=list(i.far_pairs[:L]), eval_clusters=i.eval_clusters,
                eval_radius_max=i.eval_radius_max, eval_radius_min=i.eval_radius_min,
                warning=i.warning.decode())


class _Frequency(ctypes.Structure):
    _fields_ = [("k", ctypes.c_double), ("n_rhs", ctypes.c_int32), ("bc_value", ctypes.c_void_p),
                ("admittance", ctypes.c_void_p), ("n_point_sources", ctypes.c_int32),
                ("point_sources", ctypes.c_void_p), ("n_plane_waves", ctypes.c_int32),
                ("plane_waves", ctypes.c_void_p)]


class _Options(ctypes.Structure):
    _fields_ = [("solver", ctypes.c_int32), ("precond", ctypes.c_int32), ("max_iterations", ctypes.c_int32),
                ("tolerance", ctypes.c_double), ("restart", ctypes.c_int32), ("refinement_steps", ctypes.c_int32),
                ("refinement_tolerance", ctypes.c_double), ("ilu_scanning_degree", ctypes.c_int32)]


class _Stats(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_int32 * 64), ("relative_residual", ctypes.c_double * 64),
                ("converged_all", ctypes.c_int32), ("refinement_steps", ctypes.c_int32),
                ("ms_assemble", ctypes.c_double), ("ms_solve", ctypes.c_double), ("ms_field", ctypes.c_double),
                ("ilu_nonzeros", ctypes.c_int64), ("ilu_levels", ctypes.c_int32 * 2)]


# every entry point include/ncg.h declares
EXPORTS = ("ncg_last_error", "ncg_device_count", "ncg_fmm_plan", "ncg_problem_create", "ncg_problem_destroy",
           "ncg_assemble", "ncg_solve", "ncg_problem_fmm_info", "ncg_get_solution", "ncg_field", "ncg_get_system",
           "ncg_apply", "ncg_timings", "ncg_server_create", "ncg_server_add_mesh", "ncg_server_submit",
           "ncg_server_wait", "ncg_server_query", "ncg_server_stats", "ncg_server_destroy",
           "ncg_set_kernel_timing", "ncg_kernel_stats", "ncg_server_kernel_stats", "ncg_solve_group",
           "ncg_set_tuning", "ncg_get_tuning", "ncg_apply_noalloc", "ncg_autotune", "ncg_server_tuning",
           "ncg_set_graphs", "ncg_graph_stats", "ncg_server_set_kernel_timing")


class _Tuning(ctypes.Structure):
    _fields_ = [("t_waves", ctypes.c_int32), ("d_waves", ctypes.c_int32), ("s_waves", ctypes.c_int32),
                ("tuned", ctypes.c_int32), ("near_mode", ctypes.c_int32), ("mfma_splits", ctypes.c_int32),
                ("assemble_rows", ctypes.c_int32), ("near_assembly", ctypes.c_int32)]


def _tuning_dict(t):
    return {k: getattr(t, k) for k, _ in _Tuning._fields_}


class _KernelStat(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char * 48), ("launches", ctypes.c_int64), ("total_ms", ctypes.c_double),
                ("bytes", ctypes.c_double), ("flops", ctypes.c_double)]


def _kernel_dict(arr, count):
    return {arr[i].name.decode(): dict(calls=arr[i].launches, total_ms=arr[i].total_ms, bytes=arr[i].bytes,
                                       flops=arr[i].flops)
            for i in range(min(count, len(arr)))}


class _ServerConfig(ctypes.Structure):
    _fields_ = [("device", ctypes.c_int32), ("lanes", ctypes.c_int32), ("memory_budget_bytes", ctypes.c_double),
                ("problems_per_lane", ctypes.c_int32), ("kernel_timing", ctypes.c_int32), ("group", ctypes.c_int32),
                ("autotune", ctypes.c_int32), ("tuning_file", ctypes.c_char_p), ("eager", ctypes.c_int32)]


class _Job(ctypes.Structure):
    _fields_ = [("mesh", ctypes.c_int32), ("n_freq", ctypes.c_int32), ("freq", ctypes.c_void_p),
                ("precision", ctypes.c_int32), ("solver", _Options), ("solution", ctypes.c_void_p),
                ("phi", ctypes.c_void_p), ("grad", ctypes.c_void_p)]


class _JobStatus(ctypes.Structure):
    _fields_ = [("state", ctypes.c_int32), ("error_code", ctypes.c_int32), ("freq_done", ctypes.c_int32),
                ("iterations", ctypes.c_int32), ("latency_ms", ctypes.c_double), ("queue_ms", ctypes.c_double),
                ("refinement_steps_max", ctypes.c_int32), ("relative_residual_max", ctypes.c_double),
                ("error", ctypes.c_char * 256)]


class _ServerInfo(ctypes.Structure):
    _fields_ = [("jobs_done", ctypes.c_int64), ("jobs_failed", ctypes.c_int64), ("steps_done", ctypes.c_int64),
                ("solves_done", ctypes.c_int64), ("p50_ms", ctypes.c_double), ("p99_ms", ctypes.c_double),
                ("mean_ms", ctypes.c_double), ("wall_s", ctypes.c_double), ("resident_bytes", ctypes.c_int64),
                ("peak_resident_bytes", ctypes.c_int64), ("lanes", ctypes.c_int32), ("queued_steps", ctypes.c_int32),
                ("lane_ms", ctypes.c_double * 4), ("tuned_shapes", ctypes.c_int32),
                ("graph_steps", ctypes.c_int64), ("eager_steps", ctypes.c_int64)]

_lib = None


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libncg.so is not built (%s); run __graft_entry__.build()" % LIB_PATH)
        lib = ctypes.CDLL(LIB_PATH)
        lib.ncg_last_error.restype = ctypes.c_char_p
        _lib = lib
    return _lib


def _check(rc):
    if rc != 0:
        raise NcgError(rc, load().ncg_last_error().decode())


def device_count():
    c = ctypes.c_int32(0)
    load().ncg_device_count(ctypes.byref(c))
    return c.value


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data) if a is not None and a.size else None


class _GeometryArrays:
    """Keeps the arrays an _Geometry points to alive."""

    def __init__(self, nv, xyz, kind, eval_xyz=None, h=1.0, interior=0, node_index=None, group=None,
                 method_bem=0, min_terms=0, cluster_edge=0.0):
        self.nv = np.ascontiguousarray(nv, np.int32)
        n = len(self.nv)
        x = np.zeros((n, 4, 3))
        xyz = np.asarray(xyz, np.float64)
        x[:, :xyz.shape[1], :] = xyz
        self.xyz = np.ascontiguousarray(x)
        self.kind = np.ascontiguousarray(kind if kind is not None else np.zeros(n), np.int32)
        self.eval = np.ascontiguousarray(eval_xyz if eval_xyz is not None else np.zeros((0, 3)), np.float64)
        self.node = None
        if node_index is not None:
            ni = np.full((n, 4), -1, np.int32)
            node_index = np.asarray(node_index)
            ni[:, :node_index.shape[1]] = node_index
            self.node = np.ascontiguousarray(ni)
        self.group = None if group is None else np.ascontiguousarray(group, np.int32)
        self.n, self.m = n, len(self.eval)
        self.g = _Geometry(n, self.nv.ctypes.data, self.xyz.ctypes.data, self.kind.ctypes.data, self.m,
                           self.eval.ctypes.data if self.m else None, float(h), int(interior),
                           self.node.ctypes.data if self.node is not None else None,
                           self.group.ctypes.data if self.group is not None else None,
                           _Method(int(method_bem), int(min_terms), float(cluster_edge)))


def fmm_plan(nv, xyz, k, eval_xyz=None, node_index=None, method_bem=4, min_terms=0, cluster_edge=0.0):
    """NumCalc's method and FMM plan at wave number k (host only, no device)."""
    ga = _GeometryArrays(nv, xyz, None, eval_xyz, node_index=node_index, method_bem=method_bem,
                         min_terms=min_terms, cluster_edge=cluster_edge)
    info = _FmmInfo()
    _check(load().ncg_fmm_plan(ctypes.byref(ga.g), ctypes.c_double(k), ctypes.byref(info)))
    return _info_dict(info)


class Problem:
    """One boundary-element problem resident on a GPU (geometry, near-pair table)."""

    def __init__(self, nv, xyz, kind, eval_xyz=None, h=1.0, interior=0, device=0, node_index=None, group=None,
                 method_bem=0, min_terms=0, cluster_edge=0.0):
        self.lib = load()
        self._ga = _GeometryArrays(nv, xyz, kind, eval_xyz, h, interior, node_index, group, method_bem, min_terms,
                                   cluster_edge)
        self.n, self.m = self._ga.n, self._ga.m
        self._h = ctypes.c_void_p()
        _check(self.lib.ncg_problem_create(ctypes.byref(self._ga.g), int(device), ctypes.byref(self._h)))
        self.nrhs = 0

    def fmm_info(self):
        info = _FmmInfo()
        _check(self.lib.ncg_problem_fmm_info(self._h, ctypes.byref(info)))
        return _info_dict(info)

    def close(self):
        if self._h:
            self.lib.ncg_problem_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def assemble(self, k, bc=None, adm=None, point_sources=(), plane_waves=(), precision=FP64, nrhs=None):
        """bc: (nrhs, n) complex prescribed values; sources: rows (x, y, z, re, im, rhs)."""
        if bc is None:
            bc = np.zeros((nrhs or 1, self.n), np.complex128)
        bc = np.ascontiguousarray(np.atleast_2d(bc), np.complex128)
        self.nrhs = bc.shape[0]
        self._bc = bc
        self._adm = np.ascontiguousarray(adm if adm is not None else np.zeros(self.n), np.complex128)
        self._ps = np.ascontiguousarray(np.asarray(point_sources, np.float64).reshape(-1, 6))
        self._pw = np.ascontiguousarray(np.asarray(plane_waves, np.float64).reshape(-1, 6))
        f = _Frequency(float(k), self.nrhs, bc.ctypes.data, self._adm.ctypes.data, len(self._ps), _ptr(self._ps),
                       len(self._pw), _ptr(self._pw))
        _check(self.lib.ncg_assemble(self._h, ctypes.byref(f), int(precision)))

    def solve(self, solver=CGS, precond=NONE, max_iterations=250, tolerance=1e-9, restart=50,
              refinement_steps=8, refinement_tolerance=1e-6, allow_noconv=False, ilu_scanning_degree=0):
        o = _Options(solver, precond, max_iterations, tolerance, restart, refinement_steps, refinement_tolerance,
                     ilu_scanning_degree)
        st = _Stats()
        rc = self.lib.ncg_solve(self._h, ctypes.byref(o), ctypes.byref(st))
        if rc != 0 and not (allow_noconv and rc == -6):
            _check(rc)
        return dict(iterations=list(st.iterations[:self.nrhs]),
                    relative_residual=list(st.relative_residual[:self.nrhs]),
                    converged=bool(st.converged_all), refinement_steps=st.refinement_steps,
                    ms_assemble=st.ms_assemble, ms_solve=st.ms_solve,
                    ilu_nonzeros=st.ilu_nonzeros, ilu_levels=tuple(st.ilu_levels))

    def solution(self):
        x = np.zeros((self.nrhs, self.n), np.complex128)
        _check(self.lib.ncg_get_solution(self._h, ctypes.c_void_p(x.ctypes.data)))
        return x

    def field(self, gradient=True):
        phi = np.zeros((self.nrhs, self.m), np.complex128)
        grad = np.zeros((self.nrhs, self.m, 3), np.complex128) if gradient else None
        _check(self.lib.ncg_field(self._h, ctypes.c_void_p(phi.ctypes.data), _ptr(grad) if gradient else None))
        return phi, grad

    def system(self):
        A = np.zeros((self.n, self.n), np.complex128)
        b = np.zeros((self.nrhs, self.n), np.complex128)
        _check(self.lib.ncg_get_system(self._h, ctypes.c_void_p(A.ctypes.data), ctypes.c_void_p(b.ctypes.data)))
        return A, b

    def apply(self, x):
        x = np.ascontiguousarray(np.atleast_2d(x), np.complex128)
        y = np.zeros_like(x)
        _check(self.lib.ncg_apply(self._h, ctypes.c_void_p(x.ctypes.data), ctypes.c_void_p(y.ctypes.data), len(x)))
        return y

    def kernel_timing(self, on=True):
        _check(self.lib.ncg_set_kernel_timing(self._h, int(on)))

    def graphs(self, on=True):
        """ncg_set_graphs: GMRES Arnoldi steps replayed as captured HIP graphs (default) or
        launched kernel by kernel"""
        _check(self.lib.ncg_set_graphs(self._h, int(on)))

    def graph_stats(self):
        """ncg_graph_stats: Arnoldi-step graphs captured, steps replayed, steps launched one by one"""
        out = (ctypes.c_int64 * 4)()
        _check(self.lib.ncg_graph_stats(self._h, out))
        return dict(zip(("captured", "replayed", "eager", "cached"), list(out)))

    def kernel_stats(self, reset=False):
        arr = (_KernelStat * 64)()
        cnt = ctypes.c_int32()
        _check(self.lib.ncg_kernel_stats(self._h, arr, 64, ctypes.byref(cnt), int(reset)))
        return _kernel_dict(arr, cnt.value)

    def tuning(self):
        t = _Tuning()
        _check(self.lib.ncg_get_tuning(self._h, ctypes.byref(t)))
        return _tuning_dict(t)

    def set_tuning(self, t_waves=8, d_waves=8, s_waves=8, near_mode=0, mfma_splits=0, assemble_rows=0, near_assembly=0):
        t = _Tuning(int(t_waves), int(d_waves), int(s_waves), 0, int(near_mode), int(mfma_splits), int(assemble_rows),
                    int(near_assembly))
        _check(self.lib.ncg_set_tuning(self._h, ctypes.byref(t)))

    def timings(self, reset=False):
        t = (ctypes.c_double * 8)()
        _check(self.lib.ncg_timings(self._h, t, int(reset)))
        keys = ("ms_assemble", "ms_solve", "ms_field", "far_kernel_ms", "matvec_kernel_ms", "matvec_calls",
                "quadrature_evaluations", "operator_bytes")
        return dict(zip(keys, list(t)))


def solve_group(problems, solver=GMRES, precond=BLOCK, max_iterations=500, tolerance=1e-9, restart=50,
                refinement_steps=8, refinement_tolerance=1e-6, allow_noconv=False):
    """ncg_solve_group: solve several assembled problems (equal n_rhs and precision) in
    lockstep, sharing every Krylov launch; returns one stats dict per problem."""
    lib = load()
    arr = (ctypes.c_void_p * len(problems))(*[p._h.value for p in problems])
    o = _Options(solver, precond, max_iterations, tolerance, restart, refinement_steps, refinement_tolerance, 0)
    st = (_Stats * len(problems))()
    rc = lib.ncg_solve_group(arr, len(problems), ctypes.byref(o), st)
    if rc != 0 and not (allow_noconv and rc == -6):
        _check(rc)
    return [dict(iterations=list(st[i].iterations[:p.nrhs]), relative_residual=list(st[i].relative_residual[:p.nrhs]),
                 converged=bool(st[i].converged_all), refinement_steps=st[i].refinement_steps,
                 ms_solve=st[i].ms_solve) for i, p in enumerate(problems)]


def autotune(problems, reps=3):
    """ncg_autotune: time the candidate launch shapes on these assembled problems, set the
    fastest on each, and return it."""
    lib = load()
    arr = (ctypes.c_void_p * len(problems))(*[p._h.value for p in problems])
    t = _Tuning()
    _check(lib.ncg_autotune(arr, len(problems), int(reps), ctypes.byref(t)))
    return _tuning_dict(t)


def _status_dict(st):
    return dict(state=("queued", "running", "done", "failed")[st.state], error_code=st.error_code,
                freq_done=st.freq_done, iterations=st.iterations, latency_ms=st.latency_ms, queue_ms=st.queue_ms,
                refinement_steps_max=st.refinement_steps_max, relative_residual_max=st.relative_residual_max,
                error=st.error.decode())


class ServerJob:
    """A submitted job: its id and the output arrays the server writes into."""

    def __init__(self, server, job_id, solution, phi, grad):
        self.server, self.id = server, job_id
        self.solution, self.phi, self.grad = solution, phi, grad

    def wait(self, timeout=-1.0):
        return self.server.wait(self, timeout)


class Server:
    """Serving runtime (ncg_server_*): jobs of (mesh, frequencies, load cases) on one GPU
    with continuous batching over `lanes` concurrent steps and resident problems."""

    def __init__(self, device=0, lanes=4, memory_budget=0.0, problems_per_lane=2, kernel_timing=False, group=0,
                 autotune=False, tuning_file=None, graphs=True):
        self.lib = load()
        self._h = ctypes.c_void_p()
        cfg = _ServerConfig(int(device), int(lanes), float(memory_budget), int(problems_per_lane), int(kernel_timing),
                            int(group), int(bool(autotune)), tuning_file.encode() if tuning_file else None,
                            int(not graphs))
        _check(self.lib.ncg_server_create(ctypes.byref(cfg), ctypes.byref(self._h)))
        self._meshes = {}
        # arrays the server reads or writes for each job that has not ended: they stay
        # referenced here until wait() sees the end (or the server is closed)
        self._live = {}

    def add_mesh(self, nv, xyz, kind, eval_xyz=None, h=1.0, interior=0, node_index=None, group=None, method_bem=4,
                 min_terms=0, cluster_edge=0.0):
        ga = _GeometryArrays(nv, xyz, kind, eval_xyz, h, interior, node_index, group, method_bem, min_terms,
                             cluster_edge)
        mid = ctypes.c_int32()
        _check(self.lib.ncg_server_add_mesh(self._h, ctypes.byref(ga.g), ctypes.byref(mid)))
        self._meshes[mid.value] = (ga.n, ga.m)
        return mid.value

    def submit(self, mesh, freqs, precision=MIXED, solver=GMRES, precond=BLOCK, max_iterations=500, tolerance=1e-9,
               restart=50, refinement_steps=8, refinement_tolerance=1e-6, ilu_scanning_degree=0, solution=True,
               field=True, gradient=False):
        """freqs: list of dicts with k and optionally bc ((nrhs, n) complex), adm (n complex),
        point_sources / plane_waves (rows x y z re im rhs), nrhs."""
        n, m = self._meshes[mesh]
        keep = []
        arr = (_Frequency * len(freqs))()
        nrhs = None
        for i, f in enumerate(freqs):
            nr = int(f.get("nrhs", 1 if f.get("bc") is None else np.atleast_2d(f["bc"]).shape[0]))
            nrhs = nr if nrhs is None else nrhs
            bc = f.get("bc")
            bc = None if bc is None else np.ascontiguousarray(np.atleast_2d(bc), np.complex128)
            adm = f.get("adm")
            adm = None if adm is None else np.ascontiguousarray(adm, np.complex128)
            ps = np.ascontiguousarray(np.asarray(f.get("point_sources", ()), np.float64).reshape(-1, 6))
            pw = np.ascontiguousarray(np.asarray(f.get("plane_waves", ()), np.float64).reshape(-1, 6))
            keep += [bc, adm, ps, pw]
            arr[i] = _Frequency(float(f["k"]), nr, _ptr(bc), _ptr(adm), len(ps), _ptr(ps), len(pw), _ptr(pw))
        nf = len(freqs)
        x = np.zeros((nf, nrhs, n), np.complex128) if solution else None
        phi = np.zeros((nf, nrhs, m), np.complex128) if field and m else None
        grad = np.zeros((nf, nrhs, m, 3), np.complex128) if gradient and m else None
        o = _Options(solver, precond, max_iterations, tolerance, restart, refinement_steps, refinement_tolerance,
                     ilu_scanning_degree)
        job = _Job(int(mesh), nf, ctypes.cast(arr, ctypes.c_void_p), int(precision), o, _ptr(x), _ptr(phi), _ptr(grad))
        jid = ctypes.c_int64()
        _check(self.lib.ncg_server_submit(self._h, ctypes.byref(job), ctypes.byref(jid)))
        self._live[jid.value] = (keep, arr, x, phi, grad)
        return ServerJob(self, jid.value, x, phi, grad)

    def wait(self, job, timeout=-1.0):
        st = _JobStatus()
        rc = self.lib.ncg_server_wait(self._h, ctypes.c_int64(job.id), ctypes.c_double(timeout), ctypes.byref(st))
        if rc not in (0, -5) and st.state != 3:   # a failed job reports through its status
            _check(rc)
        if st.state >= 2:
            self._live.pop(job.id, None)
        return _status_dict(st)

    def query(self, job):
        st = _JobStatus()
        _check(self.lib.ncg_server_query(self._h, ctypes.c_int64(job.id), ctypes.byref(st)))
        return _status_dict(st)

    def set_kernel_timing(self, on=True):
        """per-kernel timing of the steps started from now on (turns graph replay off while on)"""
        _check(self.lib.ncg_server_set_kernel_timing(self._h, int(on)))

    def kernel_stats(self, reset=False):
        arr = (_KernelStat * 64)()
        cnt = ctypes.c_int32()
        _check(self.lib.ncg_server_kernel_stats(self._h, arr, 64, ctypes.byref(cnt), int(reset)))
        return _kernel_dict(arr, cnt.value)

    def stats(self):
        s = _ServerInfo()
        _check(self.lib.ncg_server_stats(self._h, ctypes.byref(s)))
        d = {k: getattr(s, k) for k, _ in _ServerInfo._fields_}
        d["lane_ms"] = dict(zip(("assemble", "solve", "outputs", "waiting"), list(s.lane_ms)))
        return d

    def tunings(self):
        """[(operator shape, tuning dict)] the server has chosen (ncg_server_tuning)"""
        out = []
        for i in range(self.stats()["tuned_shapes"]):
            key = ctypes.create_string_buffer(256)
            t = _Tuning()
            _check(self.lib.ncg_server_tuning(self._h, i, key, 256, ctypes.byref(t)))
            out.append((key.value.decode(), _tuning_dict(t)))
        return out

    def close(self):
        if self._h:
            self.lib.ncg_server_destroy(self._h)   # finishes the queued work first
            self._h = ctypes.c_void_p()
        self._live.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
