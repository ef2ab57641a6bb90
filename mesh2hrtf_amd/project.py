"""NumCalc project folders: deterministic synthetic meshes, NC.inp writer, be.out reader.

The file formats follow the NumCalc contract:
  * NC.inp records and tokenizer: mesh2hrtf/NumCalc/src/NC_Input.cpp:151-347, 1596-1631
    (as emitted by mesh2hrtf/Mesh2Input/mesh2input.py:1209-1379);
  * node files ``count`` then ``id x y z``; element files ``count`` then
    ``id n1 n2 n3 [n4] prop 0 group`` (NC_Input.cpp:366-435);
  * be.out/be.N/{pBoundary,vBoundary,pEvalGrid,vEvalGrid}
    (mesh2hrtf/NumCalc/src/NC_PostProcessing.cpp:447-546, 688-749).

Every generator is seeded and pure numpy/scipy so the CPU baseline and the GPU
run see the same inputs.
"""
from __future__ import annotations

import os
from dataclasses import dataclass, field

import numpy as np

EVAL_ID0 = 300000  # first node/element id of evaluation grids (mesh2input convention)


# ----------------------------------------------------------------------------- meshes

@dataclass
class Mesh:
    nodes: np.ndarray            # (n, 3) float64
    elems: np.ndarray            # (e, 3) or (e, 4) int, 0-based rows of ``nodes``
    node_ids: np.ndarray = None  # external ids
    elem_ids: np.ndarray = None

    def __post_init__(self):
        self.nodes = np.ascontiguousarray(self.nodes, dtype=np.float64)
        self.elems = np.ascontiguousarray(self.elems, dtype=np.int64)
        if self.node_ids is None:
            self.node_ids = np.arange(len(self.nodes), dtype=np.int64)
        if self.elem_ids is None:
            self.elem_ids = np.arange(len(self.elems), dtype=np.int64)


def _icosahedron():
    t = (1.0 + 5.0 ** 0.5) / 2.0
    v = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0],
                  [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
                  [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], dtype=np.float64)
    f = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11],
                  [1, 5, 9], [5, 11, 4], [11, 10, 2], [10, 7, 6], [7, 1, 8],
                  [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9],
                  [4, 9, 5], [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]], dtype=np.int64)
    return v / np.linalg.norm(v, axis=1, keepdims=True), f


def _orient_outward(nodes, tris):
    """Make every triangle counter-clockwise seen from outside (star-shaped meshes).

    NumCalc's element normal is (x0-x2) x (x1-x2) (NC_3dFunctions.cpp:143-175), which
    points outward exactly for this ordering."""
    a, b, c = nodes[tris[:, 0]], nodes[tris[:, 1]], nodes[tris[:, 2]]
    n = np.cross(a - c, b - c)
    flip = np.einsum("ij,ij->i", n, (a + b + c) / 3.0) < 0
    tris = tris.copy()
    tris[flip, 0], tris[flip, 1] = tris[flip, 1], tris[flip, 0].copy()
    return tris


def icosphere(subdivisions: int, radius: float) -> Mesh:
    """Icosahedron refined by midpoint subdivision, projected to a sphere: 20*4^s triangles."""
    v, f = _icosahedron()
    verts = [tuple(x) for x in v]
    for _ in range(subdivisions):
        cache = {}
        nf = []

        def mid(i, j):
            key = (min(i, j), max(i, j))
            if key not in cache:
                m = (np.asarray(verts[i]) + np.asarray(verts[j])) * 0.5
                m /= np.linalg.norm(m)
                verts.append(tuple(m))
                cache[key] = len(verts) - 1
            return cache[key]

        for a, b, c in f:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [[a, ab, ca], [b, bc, ab], [c, ca, bc], [ab, bc, ca]]
        f = np.array(nf, dtype=np.int64)
    nodes = np.asarray(verts) * radius
    return Mesh(nodes, _orient_outward(nodes, f))


def geodesic_sphere(freq: int) -> Mesh:
    """Class-I geodesic unit sphere: every icosahedron face split into freq^2 triangles."""
    v, f = _icosahedron()
    index = {}
    pts = []

    def pid(p):
        key = tuple(np.round(p, 12))
        if key not in index:
            index[key] = len(pts)
            pts.append(p)
        return index[key]

    tris = []
    for a, b, c in f:
        A, B, C = v[a], v[b], v[c]
        ids = {}
        for i in range(freq + 1):
            for j in range(freq + 1 - i):
                p = A + (B - A) * (i / freq) + (C - A) * (j / freq)
                ids[i, j] = pid(p / np.linalg.norm(p))
        for i in range(freq):
            for j in range(freq - i):
                tris.append([ids[i, j], ids[i + 1, j], ids[i, j + 1]])
                if i + j < freq - 1:
                    tris.append([ids[i + 1, j], ids[i + 1, j + 1], ids[i, j + 1]])
    nodes = np.asarray(pts)
    return Mesh(nodes, _orient_outward(nodes, np.asarray(tris, dtype=np.int64)))


def _real_sph_harm_matrix(lmax, dirs):
    """Orthonormal real spherical harmonics Y_lm(dirs) for l<=lmax, shape (n, (lmax+1)^2)."""
    from scipy.special import sph_harm_y
    theta = np.arccos(np.clip(dirs[:, 2], -1.0, 1.0))
    phi = np.arctan2(dirs[:, 1], dirs[:, 0])
    cols = []
    for l in range(lmax + 1):
        for m in range(-l, l + 1):
            y = sph_harm_y(l, abs(m), theta, phi)
            if m > 0:
                cols.append(np.sqrt(2.0) * (-1) ** m * y.real)
            elif m < 0:
                cols.append(np.sqrt(2.0) * (-1) ** m * y.imag)
            else:
                cols.append(y.real)
    return np.stack(cols, axis=1)


def head_like_mesh(freq: int = 32, seed: int = 0, semi_axes=(0.075, 0.09, 0.11),
                   lmax: int = 8, rms: float = 3e-3) -> Mesh:
    """Seeded head-like closed surface: ellipsoid radially perturbed by a real SH sum.

    Coefficients c_lm ~ N(0,1) (numpy default_rng(seed)) for l<=lmax, scaled so the RMS
    radial displacement over the sphere is ``rms``; the geodesic sphere of frequency
    ``freq`` (20*freq^2 triangles) is mapped radially onto the surface."""
    g = geodesic_sphere(freq)
    u = g.nodes
    rng = np.random.default_rng(seed)
    coef = rng.standard_normal((lmax + 1) ** 2)
    coef *= rms / np.sqrt(np.sum(coef ** 2) / (4.0 * np.pi))
    a, b, c = semi_axes
    r_ell = 1.0 / np.sqrt((u[:, 0] / a) ** 2 + (u[:, 1] / b) ** 2 + (u[:, 2] / c) ** 2)
    r = r_ell + _real_sph_harm_matrix(lmax, u) @ coef
    nodes = u * r[:, None]
    return Mesh(nodes, _orient_outward(nodes, g.elems))


def graded_head_mesh(seed: int = 0, semi_axes=(0.075, 0.09, 0.11), lmax: int = 8, rms: float = 3e-3,
                     base: int = 14, bands=(0.03, 0.12, 0.135)) -> Mesh:
    """Head-like surface (the head_like_mesh surface of the same seed) meshed graded towards
    the ears: the nodes of the geodesic spheres of frequency base * 2^l (l = 3, 2, 1) whose
    great-circle distance (on the 0.09 m sphere) to the nearer of the +-y ear directions is
    below bands[0], bands[1], bands[2], plus every node of frequency base, triangulated as
    their convex hull on the unit sphere and mapped radially onto the surface.  Defaults:
    60320 triangles, edges about 1 mm within 30 mm of the ears, 2 mm to 120 mm, 4 mm to
    135 mm and 8 mm on the far sides."""
    from scipy.spatial import ConvexHull
    ears = np.array([[0.0, 1.0, 0.0], [0.0, -1.0, 0.0]])
    pts = [geodesic_sphere(base).nodes]
    for level, band in zip((3, 2, 1), bands):
        u = geodesic_sphere(base * 2 ** level).nodes
        d = 0.09 * np.arccos(np.clip(np.max(u @ ears.T, axis=1), -1.0, 1.0))
        pts.append(u[d < band])
    u = np.unique(np.round(np.concatenate(pts), 12), axis=0)
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    tris = ConvexHull(u).simplices.astype(np.int64)
    rng = np.random.default_rng(seed)
    coef = rng.standard_normal((lmax + 1) ** 2)
    coef *= rms / np.sqrt(np.sum(coef ** 2) / (4.0 * np.pi))
    a, b, c = semi_axes
    r_ell = 1.0 / np.sqrt((u[:, 0] / a) ** 2 + (u[:, 1] / b) ** 2 + (u[:, 2] / c) ** 2)
    nodes = u * (r_ell + _real_sph_harm_matrix(lmax, u) @ coef)[:, None]
    return Mesh(nodes, _orient_outward(nodes, tris))


def config3_freqs():
    """12000..23700 Hz step 300 Hz (40 frequencies)."""
    return [12000.0 + 300.0 * i for i in range(40)]


def ear_sources(mesh: Mesh, offset: float = 1e-3):
    """Two point sources ``offset`` outside the surface nodes nearest the +y and -y axes."""
    u = mesh.nodes / np.linalg.norm(mesh.nodes, axis=1, keepdims=True)
    out = []
    for s in (+1.0, -1.0):
        k = int(np.argmax(s * u[:, 1]))
        out.append(mesh.nodes[k] + offset * u[k])
    return np.asarray(out)


def fibonacci_grid(m: int, radius: float, id0: int = EVAL_ID0) -> Mesh:
    """Fibonacci lattice on a sphere, triangulated by its convex hull.

    z_k = 1 - 2(k+1/2)/M, phi_k = pi(1+sqrt 5)k; ids start at ``id0``."""
    from scipy.spatial import ConvexHull
    k = np.arange(m, dtype=np.float64)
    z = 1.0 - 2.0 * (k + 0.5) / m
    phi = np.pi * (1.0 + 5.0 ** 0.5) * k
    s = np.sqrt(1.0 - z * z)
    pts = radius * np.stack([s * np.cos(phi), s * np.sin(phi), z], axis=1)
    hull = ConvexHull(pts)
    tris = _orient_outward(pts, hull.simplices.astype(np.int64))
    return Mesh(pts, tris, node_ids=id0 + np.arange(m), elem_ids=id0 + np.arange(len(tris)))


# ----------------------------------------------------------------------------- NC.inp

@dataclass
class Boundary:
    lo: int
    hi: int
    kind: str          # VELO | PRES | ADMI | IMPE
    re: float
    im: float
    # frequency curves scaling the real and imaginary part (NC.inp CURVES ids; -1 none)
    re_curve: int = -1
    im_curve: int = -1


@dataclass
class Project:
    mesh: Mesh
    freqs: list
    eval_grid: Mesh | None = None
    boundary: list = field(default_factory=list)
    point_sources: list = field(default_factory=list)   # (x, y, z, re, im)
    plane_waves: list = field(default_factory=list)     # (dx, dy, dz, re, im)
    method_bem: int = 0          # 0 TBEM, 1 SL-FMM, 4 ML-FMM
    method_solver: int = 0       # 0 CGS, 4 Gauss
    precond: int = 0             # IPRECOND
    c: float = 343.0
    rho: float = 1.1839
    title: str = "synthetic project"
    # symmetry planes of NC.inp (src/NC_Input.cpp:459-495): per axis x, y, z the plane type
    # (0 none, 1 symmetric, 2 antisymmetric, 3 and 4 as NumCalc reads them) and its coordinate
    symmetry: tuple | None = None
    # NC.inp point-source count -1 (NumCalc's "test point source"): no point source is read
    test_point_source: bool = False
    # IFINTE (src/NC_Input.cpp:296): 1 solves the interior problem
    interior: bool = False
    # CURVES block (src/NC_Input.cpp:1090-1145): (id, [(frequency, factor), ...]) per curve;
    # the curves of each source's real and imaginary amplitude, (-1, -1) for none
    curves: list = field(default_factory=list)
    point_source_curves: list = field(default_factory=list)
    plane_wave_curves: list = field(default_factory=list)


def _fmt_freq_curve(freqs):
    lines = ["0 %d" % (len(freqs) + 1), "0.000000 0.000000e+00 0.0"]
    for i, f in enumerate(freqs):
        lines.append("%.6f %.16e 0.0" % ((i + 1) * 1e-6, f))
    return lines


def _write_mesh_files(folder, mesh, prop, group):
    os.makedirs(folder, exist_ok=True)
    with open(os.path.join(folder, "Nodes.txt"), "w") as fh:
        fh.write("%d\n" % len(mesh.nodes))
        for i, p in zip(mesh.node_ids, mesh.nodes):
            fh.write("%d %.17g %.17g %.17g\n" % (i, p[0], p[1], p[2]))
    with open(os.path.join(folder, "Elements.txt"), "w") as fh:
        fh.write("%d\n" % len(mesh.elems))
        for i, e in zip(mesh.elem_ids, mesh.elems):
            ids = " ".join(str(int(mesh.node_ids[n])) for n in e)
            fh.write("%d %s %d 0 %d\n" % (i, ids, prop, group))


def write_project(folder: str, prj: Project) -> str:
    """Write NC.inp plus ObjectMeshes/EvaluationGrids into ``folder``; returns NC.inp path."""
    os.makedirs(folder, exist_ok=True)
    _write_mesh_files(os.path.join(folder, "ObjectMesh"), prj.mesh, 0, 0)
    n_nodes = len(prj.mesh.nodes)
    n_elems = len(prj.mesh.elems)
    groups = 1
    if prj.eval_grid is not None:
        _write_mesh_files(os.path.join(folder, "EvalGrid"), prj.eval_grid, 2, 1)
        n_nodes += len(prj.eval_grid.nodes)
        n_elems += len(prj.eval_grid.elems)
        groups = 2
    nsym = 0 if prj.symmetry is None else sum(1 for t in prj.symmetry[0] if t)
    bnd = prj.boundary or [Boundary(int(prj.mesh.elem_ids[0]), int(prj.mesh.elem_ids[-1]),
                                    "VELO", 0.0, 0.0)]
    L = ["##", "## synthetic NumCalc project", "##", "Mesh2HRTF 1.0.0", "##", prj.title, "##",
         "## Controlparameter I", "0 0 0 0 7 0", "##", "## Controlparameter II",
         "1 %d 0.000001 0.00e+00 1 0 0" % len(prj.freqs), "##", "## Load Frequency Curve"]
    L += _fmt_freq_curve(prj.freqs)
    L += ["##", "## 1. Main Parameters I",
          "%d %d %d 0 %d 2 1 %d %d" % (groups, n_elems, n_nodes, nsym, prj.method_bem, prj.method_solver),
          "##", "## 2. Main Parameters II",
          "%d %d %d 0.0000e+00 %d 0 0" % (len(prj.plane_waves), -1 if prj.test_point_source else len(prj.point_sources),
                                          1 if prj.interior else 0, prj.precond),
          "##", "## 3. Main Parameters III", "0 0 0 0", "##", "## 4. Main Parameters IV",
          "%s %s 1.0 0.0e+00 0.0 e+00 0.0e+00 0.0e+00" % (repr(prj.c), repr(prj.rho)), "##",
          "NODES", "ObjectMesh/Nodes.txt"]
    if prj.eval_grid is not None:
        L.append("EvalGrid/Nodes.txt")
    L += ["##", "ELEMENTS", "ObjectMesh/Elements.txt"]
    if prj.eval_grid is not None:
        L.append("EvalGrid/Elements.txt")
    L.append("##")
    if nsym:
        L += ["SYMMETRY", "%d %d %d" % tuple(prj.symmetry[0]), "%.17g %.17g %.17g" % tuple(prj.symmetry[1]), "##"]
    L.append("BOUNDARY")
    for b in bnd:
        L.append("ELEM %d TO %d %s %.17g %d %.17g %d" % (b.lo, b.hi, b.kind, b.re, b.re_curve, b.im, b.im_curve))
    L += ["RETU", "##"]
    if prj.plane_waves:
        L.append("PLANE WAVES")
        for i, (dx, dy, dz, re, im) in enumerate(prj.plane_waves):
            cr, ci = prj.plane_wave_curves[i] if prj.plane_wave_curves else (-1, -1)
            L.append("%d %.17g %.17g %.17g %.17g %d %.17g %d" % (i, dx, dy, dz, re, cr, im, ci))
        L.append("##")
    if prj.point_sources and not prj.test_point_source:
        L.append("POINT SOURCES")
        for i, (x, y, z, re, im) in enumerate(prj.point_sources):
            cr, ci = prj.point_source_curves[i] if prj.point_source_curves else (-1, -1)
            L.append("%d %.17g %.17g %.17g %.17g %d %.17g %d" % (i, x, y, z, re, cr, im, ci))
        L.append("##")
    if prj.curves:
        L += ["CURVES", "%d %d" % (len(prj.curves), max(len(pts) for _, pts in prj.curves))]
        for cid, pts in prj.curves:
            L.append("%d %d" % (cid, len(pts)))
            L += ["%.17g %.17g" % (f, v) for f, v in pts]
        L.append("##")
    L += ["POST PROCESS", "##", "END"]
    path = os.path.join(folder, "NC.inp")
    with open(path, "w") as fh:
        fh.write("\n".join(L) + "\n")
    return path


def write_sources(folder: str, prj: Project, sources) -> list:
    """A Mesh2HRTF-style project folder: one ``NumCalc/source_<k>`` NumCalc project per point
    source (k = 1, 2, ...), each the project ``prj`` with that source alone (amplitude 1),
    as mesh2input lays out reciprocal simulations.  Returns the source folders."""
    out = []
    for k, s in enumerate(sources, start=1):
        d = os.path.join(folder, "NumCalc", "source_%d" % k)
        p = Project(**{**prj.__dict__, "point_sources": [(float(s[0]), float(s[1]), float(s[2]), 1.0, 0.0)]})
        write_project(d, p)
        out.append(d)
    return out


# ----------------------------------------------------------------------------- configs

def config1_project(freqs=None) -> Project:
    """Rigid sphere r=0.1 m (icosphere-3, 1280 triangles), point source (0,0,0.2) m,
    2030-point Fibonacci grid at 1.2 m, TBEM + CGS, 100..2000 Hz step 100 Hz."""
    if freqs is None:
        freqs = [100.0 * (i + 1) for i in range(20)]
    return Project(mesh=icosphere(3, 0.1), freqs=list(freqs),
                   eval_grid=fibonacci_grid(2030, 1.2),
                   point_sources=[(0.0, 0.0, 0.2, 1.0, 0.0)],
                   title="rigid sphere r=0.1 m icosphere-3, point source at (0,0,0.2)")


def config2_mesh(freq: int = 32) -> Mesh:
    """Seeded head-like mesh (ellipsoid 0.075/0.09/0.11 m + SH perturbation, RMS 3 mm)."""
    return head_like_mesh(freq=freq, seed=0)


def config2_freqs():
    return [200.0 * (i + 1) for i in range(80)]


def batched_problems(count: int, seed: int = 0):
    """Batched uneven dense problems: seeded ellipsoids (even index) and perturbed spheres
    (odd index), radii 0.05-0.12 m, 500-6000 triangles log-uniform, each with 8 frequencies
    uniform in 0.2-8 kHz; 16 plane-wave load cases with directions uniform on the sphere
    (normalised Gaussian 3-vectors, drawn from their own generator so they do not depend on
    ``count``).  Returns ([(mesh, freqs)], dirs)."""
    rng = np.random.default_rng(seed)
    meshes = []
    for i in range(count):
        fr = max(5, int(round(np.sqrt(np.exp(rng.uniform(np.log(500.0), np.log(6000.0))) / 20.0))))
        radius = rng.uniform(0.05, 0.12)
        if i % 2 == 0:
            axes, rms = tuple(radius * rng.uniform(0.7, 1.0, 3)), 0.0
        else:
            axes, rms = (radius, radius, radius), 0.03 * radius
        meshes.append((head_like_mesh(freq=fr, seed=1000 + i, semi_axes=axes, rms=rms),
                       list(rng.uniform(200.0, 8000.0, 8))))
    dirs = np.random.default_rng(seed + 1).standard_normal((16, 3))
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    return meshes, dirs


def plane_wave_spec(dirs, f, c: float = 343.0, rho: float = 1.1839):
    """One frequency step with one plane-wave load case per direction, amplitude 1 in NC.inp
    terms (the library takes source strengths divided by i rho omega)."""
    omega = 2 * np.pi * f
    amp = 1.0 / (1j * rho * omega)
    return dict(k=omega / c, nrhs=len(dirs),
                plane_waves=[[d[0], d[1], d[2], amp.real, amp.imag, r] for r, d in enumerate(dirs)])


def point_source_spec(src, f, c: float = 343.0, rho: float = 1.1839):
    """One frequency step with one point-source load case per source position, amplitude 1
    in NC.inp terms."""
    omega = 2 * np.pi * f
    amp = 1.0 / (1j * rho * omega)
    return dict(k=omega / c, nrhs=len(src),
                point_sources=[[s[0], s[1], s[2], amp.real, amp.imag, r] for r, s in enumerate(src)])


# ----------------------------------------------------------------------------- be.out

def read_be_file(path):
    """Parse one be.out file into (ids, values): values complex (n,) or (n, 3) for vEvalGrid."""
    with open(path) as fh:
        lines = fh.read().splitlines()
    ngroups = int(lines[1].split()[0])
    ids, vals = [], []
    k = 2
    for _ in range(ngroups):
        if k >= len(lines):
            break
        _, cnt = lines[k].split()[:2]
        k += 1
        for line in lines[k:k + int(cnt)]:
            t = line.split()
            ids.append(int(t[0]))
            x = [float(v) for v in t[1:]]
            vals.append([complex(x[i], x[i + 1]) for i in range(0, len(x), 2)])
        k += int(cnt)
    v = np.asarray(vals)
    if v.shape[1] == 1:
        v = v[:, 0]
    return np.asarray(ids), v


def rel_l2(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))
