"""Test helpers: golden fixtures, NumCalc input transforms, and the oracle binding.

The transforms restate what NumCalc does between NC.inp and the solve:
  * boundary lines -> per-element kind/value (src/NC_Input.cpp:1163-1406);
  * per-frequency scaling (NC_UpdateFreqCurves, src/NC_EquationSystem.cpp:1815-1961):
    pressure / (i rho omega h), admittance * (i h k), sources / (i rho omega h);
  * outputs p = i rho omega h phi (src/NC_PostProcessing.cpp:541-545, 744).
"""
import ctypes
import json
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def load_case(name):
    z = np.load(os.path.join(GOLDEN, "case_%s.npz" % name))   # allow_pickle stays False
    d = {k: z[k] for k in z.files}
    d["spec"] = json.loads(str(d["spec"]))
    return d


def case_names():
    return sorted(f[5:-4] for f in os.listdir(GOLDEN) if f.startswith("case_") and f.endswith(".npz"))


def element_bc(case):
    """Per-element (kind, nodal value (n,4), admittance, admittance curve ids (n,2)) as given
    in NC.inp, before scaling."""
    elems = case["elems"]
    n = len(elems)
    ext = case["elem_ids"]
    loc = {int(e): i for i, e in enumerate(ext)}
    kind = np.zeros(n, dtype=np.int32)
    val = np.zeros((n, 4), dtype=np.complex128)
    adm = np.zeros(n, dtype=np.complex128)
    acurve = np.full((n, 2), -1, dtype=np.int64)
    for line in case["spec"]["boundary"]:
        lo, hi, key, re, im = line[:5]
        curves = tuple(line[5:7]) if len(line) >= 7 else (-1, -1)
        z = complex(re, im)
        nb = {"VELO": 0, "PRES": 1, "ADMI": 2, "IMPE": 2}[key]
        if key == "IMPE":
            if re == 0 and im == 0:
                nb = 1
            else:
                z = 1.0 / z
        for e in range(lo, hi + 1):
            if e not in loc:
                continue
            i = loc[e]
            if nb == 0:
                val[i, :] = z
            elif nb == 1:
                val[i, :] = z
                kind[i] = 1
            else:
                adm[i] = z
                if curves[0] >= 0:   # NumCalc keeps curves of admittance lines only
                    acurve[i] = curves
                if kind[i] == 0:
                    kind[i] = 2
    return kind, val, adm, acurve


def curve_factor(spec, cid, f):
    """Factor of frequency curve ``cid`` at ``f``: linear between its points, its end values
    outside them (src/NC_EquationSystem.cpp:1963-2009)."""
    for c, pts in spec.get("curves", []):
        if c == cid:
            fr = np.array([p[0] for p in pts])
            fa = np.array([p[1] for p in pts])
            return float(np.interp(f, fr, fa))
    raise KeyError("curve %d" % cid)


def _scaled(spec, z, curves, f):
    """A source amplitude or admittance with its real and imaginary part scaled by curves."""
    re, im = z.real, z.imag
    if curves[0] >= 0:
        re *= curve_factor(spec, curves[0], f)
    if curves[1] >= 0:
        im *= curve_factor(spec, curves[1], f)
    return complex(re, im)


def problem_arrays(case, freq, h=1.0):
    """Arrays of the oracle's ncp_problem at one frequency (host-side scaling applied)."""
    spec = case["spec"]
    nodes, elems = case["nodes"], case["elems"]
    n = len(elems)
    nv = np.full(n, elems.shape[1], dtype=np.int32)
    xyz = np.zeros((n, 4, 3))
    xyz[:, :elems.shape[1], :] = nodes[elems]
    kind, val, adm, acurve = element_bc(case)
    adm = np.array([_scaled(spec, a, c, freq) for a, c in zip(adm, acurve)], np.complex128)
    omega = 2.0 * np.pi * freq
    k = omega / spec["c"]
    rpfact = spec["rho"] * omega * h
    val = val.copy()
    val[kind == 1] = val[kind == 1] / (1j * rpfact)
    adm = np.where(kind == 2, adm * (1j * h * k), 0.0)
    psc = spec.get("point_source_curves") or [(-1, -1)] * len(spec["point_sources"])
    pwc = spec.get("plane_wave_curves") or [(-1, -1)] * len(spec["plane_waves"])
    ps = np.array([[s[0], s[1], s[2], *(lambda z: (z.real, z.imag))(
        _scaled(spec, complex(s[3], s[4]), c, freq) / (1j * rpfact))]
                   for s, c in zip(spec["point_sources"], psc)]).reshape(-1, 5)
    pw = []
    for w, c in zip(spec["plane_waves"], pwc):
        d = np.asarray(w[:3], dtype=np.float64)
        d = d / np.sqrt(d @ d)
        z = _scaled(spec, complex(w[3], w[4]), c, freq) / (1j * rpfact)
        pw.append([*d, z.real, z.imag])
    pw = np.asarray(pw).reshape(-1, 5)
    # NumCalc adds incident waves to the right-hand side only when the plane-wave count plus
    # the point-source count (-1 for its "test point source") is positive
    # (src/NC_EquationSystem.cpp:247), to the evaluation grid whenever either is nonzero
    nps = -1 if spec.get("test_point_source") else len(spec["point_sources"])
    return dict(n=n, nv=nv, xyz=xyz, kind=kind, bc=val, adm=adm, ps=ps, pw=pw,
                k=k, h=h, rpfact=rpfact, rhs_waves=len(pw) + nps > 0, interior=int(bool(spec.get("interior"))))


def rhs_arrays(pa):
    """The problem arrays the right-hand side is assembled from."""
    if pa.get("rhs_waves", True):
        return pa
    return dict(pa, pw=pa["pw"][:0], ps=pa["ps"][:0])


def unfold_symmetry(pa, spec):
    """Problem arrays of the mesh unfolded across the NC.inp symmetry planes.

    NumCalc integrates over the mirror images of every element, with the factor -1 per
    antisymmetric plane (NC_ReflectElement src/NC_3dFunctions.cpp:357-452), and mirrors
    the sources (src/NC_EquationSystem.cpp:1735-1810; plane-wave phase taken relative to
    the planes).  The same problem with the images as elements of their own: image r of
    element q is element r n + q (nodes mirrored; nodes 0<->1 and, for quads, 2<->3
    swapped after an odd number of mirrors), its prescribed velocity or pressure times the
    image's sign; its solution restricted to the first n elements is NumCalc's.
    """
    sym = spec.get("symmetry")
    if not sym:
        return pa
    types = [{0: 0, 1: 1, 2: -1, 3: 2, 4: -2}[int(t)] for t in sym[0]]
    c = np.asarray(sym[1], np.float64)
    dims = [d for d in range(3) if types[d]]
    cs = np.zeros(3)
    cs[dims] = c[dims]
    xyz, nv, kind, bc, adm, ps, pw = [], [], [], [], [], [], []
    for r in range(1 << len(dims)):
        refl = [dims[b] for b in range(len(dims)) if r >> b & 1]
        sgn = float(np.prod([-1.0 if types[d] < 0 else 1.0 for d in refl]))
        x = pa["xyz"].copy()
        for d in refl:
            x[:, :, d] = 2.0 * c[d] - x[:, :, d]
        if len(refl) % 2:
            x[:, [0, 1]] = x[:, [1, 0]]
            q4 = np.where(pa["nv"] == 4)[0]
            x[np.ix_(q4, [2, 3])] = x[np.ix_(q4, [3, 2])]
        xyz.append(x)
        nv.append(pa["nv"])
        kind.append(pa["kind"])
        bc.append(pa["bc"] * sgn)
        adm.append(pa["adm"])
        for s in pa["ps"]:
            q = s.copy()
            for d in refl:
                q[d] = 2.0 * c[d] - q[d]
            q[3:5] *= sgn
            ps.append(q)
        for s in pa["pw"]:
            q = s.copy()
            for d in refl:
                q[d] = -q[d]
            z = complex(q[3], q[4]) * sgn * np.exp(-1j * pa["h"] * pa["k"] * (q[:3] @ cs))
            q[3:5] = z.real, z.imag
            pw.append(q)
    out = dict(pa)
    out.update(n=pa["n"] << len(dims), xyz=np.concatenate(xyz), nv=np.concatenate(nv), kind=np.concatenate(kind),
               bc=np.concatenate(bc), adm=np.concatenate(adm), ps=np.asarray(ps).reshape(-1, 5),
               pw=np.asarray(pw).reshape(-1, 5))
    return out


def surface_outputs(pa, sol):
    """pBoundary and vBoundary values from the solved unknowns."""
    kind = pa["kind"]
    tau = -1.0 if pa.get("interior") else 1.0   # src/NC_3dFunctions.cpp:334
    phi = np.where(kind == 1, pa["bc"][:, 0], sol)
    v = np.where(kind == 1, sol, pa["bc"][:, 0] - np.where(kind == 2, tau * sol * pa["adm"], 0.0))
    return 1j * pa["rpfact"] * phi, v


# ----------------------------------------------------------------------------- oracle

class _Problem(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int), ("nv", ctypes.c_void_p), ("xyz", ctypes.c_void_p),
                ("bc_kind", ctypes.c_void_p), ("bc", ctypes.c_void_p), ("adm", ctypes.c_void_p),
                ("n_ps", ctypes.c_int), ("ps", ctypes.c_void_p), ("n_pw", ctypes.c_int),
                ("pw", ctypes.c_void_p), ("k", ctypes.c_double), ("h", ctypes.c_double),
                ("interior", ctypes.c_int)]


class Oracle:
    """ctypes binding of oracle/libncport.so (the C restatement); test use only."""

    def __init__(self):
        path = os.path.join(ROOT, "oracle", "libncport.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle not built: make -C oracle port")
        self.lib = ctypes.CDLL(path)

    def _prob(self, pa):
        keep = dict(nv=np.ascontiguousarray(pa["nv"], np.int32),
                    xyz=np.ascontiguousarray(pa["xyz"], np.float64),
                    kind=np.ascontiguousarray(pa["kind"], np.int32),
                    bc=np.ascontiguousarray(pa["bc"], np.complex128),
                    adm=np.ascontiguousarray(pa["adm"], np.complex128),
                    ps=np.ascontiguousarray(pa["ps"], np.float64),
                    pw=np.ascontiguousarray(pa["pw"], np.float64))
        p = _Problem(pa["n"], keep["nv"].ctypes.data, keep["xyz"].ctypes.data,
                     keep["kind"].ctypes.data, keep["bc"].ctypes.data, keep["adm"].ctypes.data,
                     len(keep["ps"]), keep["ps"].ctypes.data, len(keep["pw"]), keep["pw"].ctypes.data,
                     pa["k"], pa["h"], int(pa.get("interior", 0)))
        return p, keep

    def assemble(self, pa):
        p, keep = self._prob(pa)
        n = pa["n"]
        A = np.zeros((n, n), np.complex128)
        b = np.zeros(n, np.complex128)
        rc = self.lib.ncp_assemble(ctypes.byref(p), ctypes.c_void_p(A.ctypes.data),
                                   ctypes.c_void_p(b.ctypes.data))
        if rc:
            raise RuntimeError("ncp_assemble failed")
        return A, b

    def integrals(self, pa, i, j):
        p, keep = self._prob(pa)
        out = np.zeros(5, np.complex128)
        rc = self.lib.ncp_element_integrals(ctypes.byref(p), i, j, ctypes.c_void_p(out.ctypes.data))
        if rc:
            raise RuntimeError("ncp_element_integrals failed")
        return out

    def subelements(self, pa, j, x):
        p, keep = self._prob(pa)
        s0, t0, fa = (np.zeros(220) for _ in range(3))
        order = np.zeros(220, np.int32)
        x = np.ascontiguousarray(x, np.float64)
        n = self.lib.ncp_subelements(ctypes.byref(p), j, ctypes.c_void_p(x.ctypes.data),
                                     ctypes.c_void_p(s0.ctypes.data), ctypes.c_void_p(t0.ctypes.data),
                                     ctypes.c_void_p(fa.ctypes.data), ctypes.c_void_p(order.ctypes.data))
        return s0[:n], t0[:n], fa[:n], order[:n]

    def cgs(self, A, b, precond=2, maxit=250, tol=1e-9):
        n = len(b)
        A = np.ascontiguousarray(A, np.complex128)
        b = np.ascontiguousarray(b, np.complex128)
        x = np.zeros(n, np.complex128)
        err = ctypes.c_double(0)
        it = self.lib.ncp_cgs(n, ctypes.c_void_p(A.ctypes.data), ctypes.c_void_p(b.ctypes.data),
                              ctypes.c_void_p(x.ctypes.data), precond, maxit, ctypes.c_double(tol),
                              ctypes.byref(err))
        return x, it, err.value

    def lu_solve(self, A, b):
        A = np.array(A, np.complex128, order="C")
        b = np.array(b, np.complex128)
        self.lib.ncp_lu_solve(len(b), ctypes.c_void_p(A.ctypes.data), ctypes.c_void_p(b.ctypes.data))
        return b

    def field(self, pa, sol, pts):
        p, keep = self._prob(pa)
        pts = np.ascontiguousarray(pts, np.float64)
        sol = np.ascontiguousarray(sol, np.complex128)
        m = len(pts)
        phi = np.zeros(m, np.complex128)
        grad = np.zeros((m, 3), np.complex128)
        rc = self.lib.ncp_field(ctypes.byref(p), ctypes.c_void_p(sol.ctypes.data), m,
                                ctypes.c_void_p(pts.ctypes.data), ctypes.c_void_p(phi.ctypes.data),
                                ctypes.c_void_p(grad.ctypes.data))
        if rc:
            raise RuntimeError("ncp_field failed")
        return phi, grad

    def geometry(self, pa):
        p, keep = self._prob(pa)
        n = pa["n"]
        c, nn, a = np.zeros((n, 3)), np.zeros((n, 3)), np.zeros(n)
        self.lib.ncp_element_geometry(ctypes.byref(p), ctypes.c_void_p(c.ctypes.data),
                                      ctypes.c_void_p(nn.ctypes.data), ctypes.c_void_p(a.ctypes.data))
        return c, nn, a


def rel_l2(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


TABLE_KEYS = ("Traditional BEM", "Single level fast multipole BEM", "Direct multilevel fast multipole BEM",
              "Number of equations", "Number of entries", "Nonzeros", "    Nonzeros", "Sum of nonzeros")
TABLE_HEADS = ("Informations about clusters", "Informations about the FMM expansion",
               "Informations about the canceled interacting cluster pairs")


def nc_tables(nout):
    """Per frequency step of an NC.out text: the lines describing the analysis
    (method, cancelled-pair, cluster and expansion tables, nonzero counts),
    whitespace-normalised."""
    steps = nout.split(">> S T E P   N U M B E R   A N D   F R E Q U E N C Y <<")[1:]
    res = []
    for blk in steps:
        lines, table = [], False
        for l in blk.splitlines():
            if l.startswith(TABLE_HEADS):
                table = True
                lines.append(l.strip())
                continue
            if table and (l.startswith("   ") or l.startswith(" resu. nets")) and not l.startswith("    Nonzeros"):
                lines.append(" ".join(l.split()))
                continue
            table = False
            if l.startswith(TABLE_KEYS):
                lines.append(" ".join(l.split()))
        res.append(lines)
    return res


def same_table_line(a, b, rtol=2e-5):
    """Two normalised NC.out lines agree: words equal, numbers to rtol (the
    printed 6 significant digits of radii may round differently)."""
    ta, tb = a.split(), b.split()
    if len(ta) != len(tb):
        return False
    for x, y in zip(ta, tb):
        if x == y:
            continue
        try:
            fx, fy = float(x), float(y)
        except ValueError:
            return False
        if abs(fx - fy) > rtol * max(abs(fx), abs(fy)):
            return False
    return True


# ---- the head-like configuration at its full size (fixture case_c2_head.npz) ------------
C2_FREQS = [1000.0, 4600.0, 9800.0, 15800.0]


def mesh_digest(mesh):
    """sha256 of a mesh's node coordinates and element rows (pins the seeded generator)."""
    import hashlib
    h = hashlib.sha256(np.ascontiguousarray(mesh.nodes).tobytes())
    h.update(np.ascontiguousarray(mesh.elems, np.int64).tobytes())
    return h.hexdigest()


def c2_project():
    """The 20480-element seeded head-like mesh, its +y ear point source (amplitude 1), the
    1730-point grid at 1.2 m and four frequencies of the 200-16000 Hz sweep, NumCalc
    methodBEM 4 (SL-FMM at 1 kHz, ML-FMM above), ILU-CGS."""
    from mesh2hrtf_amd import project as P
    mesh = P.config2_mesh(32)
    grid = P.fibonacci_grid(1730, 1.2)
    src = P.ear_sources(mesh)[0]
    return P.Project(mesh=mesh, freqs=list(C2_FREQS), eval_grid=grid, method_bem=4,
                     point_sources=[(src[0], src[1], src[2], 1.0, 0.0)],
                     title="head-like mesh nu=32 (20480 triangles), +y ear source")
