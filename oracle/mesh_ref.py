"""Oracle restatement of the reference's sphere-mesh pipeline and mesh loader.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg; the product (mke_amd/) never imports it.

* ``sphere_mesh_text(k)`` restates ``mke_mesh_sphere --type full --coord local
  --iter k`` (utils/icosahedron.cpp:481-624): icosahedron
  (utils/mesh_builder.cpp:52-113), projection to the unit sphere
  (mesh_builder.cpp:115-125, 210-217), k midpoint subdivisions
  (mesh_builder.cpp:143-208), south-pole boundary (icosahedron.cpp:407-421),
  Cuthill-McKee reordering without reversal (mesh_builder.cpp:316-402,
  536-585) and the text writer with four chart coordinates per point and the
  zone rule (icosahedron.cpp:248-405), including its quirk of testing v2 twice
  and never v3 (icosahedron.cpp:380-393).
* ``load_text(text)`` restates ``Mesh::load`` (src/mesh.cpp:323-519): '#predef'
  comment block, points, 1-based triangles with optional zone, boundary list;
  ``adj`` in triangle-id order (mesh.cpp:312-320); ``inner``/``outer``/``p2io``
  (mesh.cpp:487-507).

Python floats are IEEE doubles, ``math`` calls the platform libm, and '%.16f'
formatting / float() parsing are correctly rounded like glibc's printf/strtod,
so the text is byte-identical to what a C++ restatement built with
-ffp-contract=off produces on the same machine.
"""
import math

import numpy as np

_T = (1.0 + math.sqrt(5.0)) / 2.0

# utils/mesh_builder.cpp:56-103 (data that must match the reference exactly)
_ICO_POINTS = [
    (_T, 1.0, 0.0), (-_T, 1.0, 0.0), (_T, -1.0, 0.0),
    (-_T, -1.0, 0.0), (1.0, 0.0, _T), (1.0, 0.0, -_T),
    (-1.0, 0.0, _T), (-1.0, 0.0, -_T), (0.0, _T, 1.0),
    (0.0, -_T, 1.0), (0.0, _T, -1.0), (0.0, -_T, -1.0),
]
_ICO_TRIS = [
    (0, 8, 4), (1, 10, 7), (2, 9, 11), (7, 3, 1),
    (0, 5, 10), (3, 9, 6), (3, 11, 9), (8, 6, 4),
    (2, 4, 9), (3, 7, 11), (4, 2, 0), (9, 4, 6),
    (2, 11, 5), (0, 10, 8), (5, 0, 2), (10, 5, 7),
    (1, 6, 8), (1, 8, 10), (6, 1, 3), (11, 7, 5),
]


def _project(p):
    """sphere_orto_projector, mesh_builder.cpp:115-125."""
    x, y, z = p
    l = math.sqrt(x * x + y * y + z * z)
    return (x / l, y / l, z / l)


def _divide(tris, pts):
    """One midpoint subdivision, mesh_builder.cpp:143-201."""
    already = {}
    new_pts = list(pts)
    new_tris = []

    def mid(u, v):
        key = (u, v)
        got = already.get(key)
        if got is not None:
            return got
        a, b = pts[u], pts[v]
        m = _project(((a[0] + b[0]) / 2, (a[1] + b[1]) / 2, (a[2] + b[2]) / 2))
        idx = len(new_pts)
        new_pts.append(m)
        already[(u, v)] = idx
        already[(v, u)] = idx
        return idx

    for v1, v2, v3 in tris:
        p1 = mid(v1, v2)
        p2 = mid(v1, v3)
        p3 = mid(v2, v3)
        new_tris.append((v1, p1, p2))
        new_tris.append((v3, p2, p3))
        new_tris.append((v2, p3, p1))
        new_tris.append((p1, p2, p3))
    return new_tris, new_pts


def _rcm_order(tris, n_points, boundary):
    """Graph::rcm + Mesh::full_order, mesh_builder.cpp:316-402, 536-585.

    Returns old point id -> new point id."""
    bset = set(boundary)
    inner = [i for i in range(n_points) if i not in bset]
    p2io = [0] * n_points
    for i, p in enumerate(inner):
        p2io[p] = i
    ni = 0
    for i in range(n_points):
        if i in bset:
            p2io[i] = ni
            ni += 1
    adj = [[] for _ in range(n_points)]
    for tid, t in enumerate(tris):
        for v in t:
            adj[v].append(tid)
    rs = len(inner)
    data = [set() for _ in range(rs)]
    for i in range(rs):
        p = inner[i]
        for tid in adj[p]:
            for v in tris[tid]:
                if v not in bset:
                    data[i].add(p2io[v])
    deg = [len(d) for d in data]
    # neighbours ordered by degree, ties by node id (multimap insertion order)
    srt = [sorted(d, key=lambda nb: (deg[nb], nb)) for d in data]
    order = [-1] * rs
    k = 0
    start = 0
    while True:
        queue = [start]
        head = 0
        while head < len(queue):
            v = queue[head]
            head += 1
            if order[v] != -1:
                continue
            order[v] = k
            k += 1
            queue.extend(nb for nb in srt[v] if order[nb] < 0)
        rest = [i for i in range(rs) if order[i] == -1]
        if not rest:
            break
        start = rest[0]
    full = [0] * n_points
    outer_next = rs
    for i in range(n_points):
        if i not in bset:
            full[i] = order[p2io[i]]
        else:
            full[i] = outer_next
            outer_next += 1
    return full


def _u(x, y, z):
    return math.asin(z)


def _v(x, y, z):
    r = math.sqrt(x * x + y * y)
    if r < 1e-15:
        return 0.0
    v1 = math.acos(x / r)
    if y < 0:
        v1 = 2.0 * math.pi - v1
    return v1


_HEADER = (
    "# zone1 ; zone2 ; zone 3; ... \n"
    "# u = theta  [-pi/2, pi/2]\n"
    "# v = lambda [0, 2pi]\n"
    "#predef: zone 1 stdsphere1\n"
    "#predef: zone 2 stdsphere2\n"
    "# zone 1\n# x = cos(u)*cos(v) \n# y = cos(u)*sin(v) \n# z = sin(u) \n"
    "# zone 2\n# x = -cos(u)*cos(v) \n# y = -cos(u)*sin(v) \n# z = sin(u) \n"
    "#predef: zone 3 stdsphere3\n"
    "#predef: zone 4 stdsphere4\n"
    "# zone 3\n# x = cos(u)*cos(v) \n# y = cos(u)*sin(v) \n# z = -sin(u) \n"
    "# zone 4\n# x = cos(u)*cos(v) \n# y = -cos(u)*sin(v) \n# z = sin(u) \n"
)


def sphere_mesh_text(level):
    """Text of ``mke_mesh_sphere --type full --coord local --iter level``."""
    pts = [_project(p) for p in _ICO_POINTS]
    tris = list(_ICO_TRIS)
    for _ in range(level):
        tris, pts = _divide(tris, pts)
    boundary = [i for i, p in enumerate(pts) if abs(p[2] + 1.0) < 1e-14]
    order = _rcm_order(tris, len(pts), boundary)
    new_pts = [None] * len(pts)
    for i, p in enumerate(pts):
        new_pts[order[i]] = p
    pts = new_pts
    tris = [(order[a], order[b], order[c]) for a, b, c in tris]

    out = ["#cmd:mke_mesh_sphere --type full --coord local --iter %d \n" % level,
           "#build:$Id$\n",
           "#points:%d\n" % len(pts), "#triangles:%d\n" % len(tris), _HEADER]
    for x, y, z in pts:
        line = ["%.16f %.16f %.16f" % (x, y, z)]
        for (a, b, c) in ((x, y, z), (-x, -y, z), (x, -z, y), (x, z, -y)):
            line.append(" ; %.16f %.16f " % (_u(a, b, c), _v(a, b, c)))
        line.append("\n")
        out.append("".join(line))
    out.append("# triangles %d \n" % len(tris))
    half_pi = math.pi / 2
    for a, b, c in tris:
        s = "%d %d %d " % (a + 1, b + 1, c + 1)
        (x1, y1, z1), (x2, y2, z2), (x3, y3, z3) = pts[a], pts[b], pts[c]
        v1 = _v(x1, y1, z1)
        v2 = _v(x2, y2, z2)
        if z1 > 0.9 or z2 > 0.9 or z3 > 0.9:
            s += " ; 4 "
        elif z1 < -0.9 or z2 < -0.9 or z3 < -0.9:
            s += " ; 3 "
        elif (v1 < half_pi or v2 < half_pi or v2 < half_pi
              or v1 > 1.5 * math.pi or v2 > 1.5 * math.pi or v2 > 1.5 * math.pi):
            s += " ; 2 "
        out.append(s + "\n")
    out.append("# boundary \n")
    for i, p in enumerate(pts):
        if abs(p[2] + 1.0) < 1e-14:
            out.append("%d \n" % (i + 1))
    return "".join(out)


def load_text(text):
    """Mesh::load (src/mesh.cpp:323-519) -> dict of numpy arrays."""
    lines = text.split("\n")
    pos = 0
    n_lines = len(lines)
    # comment block (the '#predef' lines only name the zone charts)
    while pos < n_lines and lines[pos].startswith("#"):
        pos += 1
    pr, charts = [], []
    while pos < n_lines and lines[pos] and not lines[pos].startswith("#"):
        toks = lines[pos].split(";")
        first = [float(v) for v in toks[0].split()]
        pr.append(first[:3])
        charts.append([[float(v) for v in t.split()[:2]] for t in toks[1:] if t.split()])
        pos += 1
    pos += 1  # '# triangles' line
    tri, zone = [], []
    while pos < n_lines and lines[pos] and not lines[pos].startswith("#"):
        parts = lines[pos].split(";")
        a, b, c = (int(v) - 1 for v in parts[0].split()[:3])
        z = int(parts[1].split()[0]) if len(parts) > 1 and parts[1].split() else 1
        tri.append((a, b, c))
        zone.append(z - 1)
        pos += 1
    pos += 1  # '# boundary' line
    n_points = len(pr)
    is_b = np.zeros(n_points, dtype=np.int32)
    while pos < n_lines and lines[pos].strip() and not lines[pos].startswith("#"):
        is_b[int(lines[pos].split()[0]) - 1] = 1
        pos += 1
    n_zones = max(len(c) for c in charts)
    coords = np.zeros((n_points, n_zones, 2))
    for i, c in enumerate(charts):
        coords[i, :len(c), :] = c
    tri = np.asarray(tri, dtype=np.int32)
    counts = np.bincount(tri.ravel(), minlength=n_points)
    adj_ptr = np.zeros(n_points + 1, dtype=np.int32)
    adj_ptr[1:] = np.cumsum(counts)
    adj = np.zeros(adj_ptr[-1], dtype=np.int32)
    fill = adj_ptr[:-1].copy()
    for t, (a, b, c) in enumerate(tri):
        for v in (a, b, c):
            adj[fill[v]] = t
            fill[v] += 1
    inner = np.nonzero(is_b == 0)[0].astype(np.int32)
    outer = np.nonzero(is_b == 1)[0].astype(np.int32)
    p2io = np.zeros(n_points, dtype=np.int32)
    p2io[inner] = np.arange(len(inner), dtype=np.int32)
    p2io[outer] = np.arange(len(outer), dtype=np.int32)
    return dict(pr=np.asarray(pr), coords=coords, tri=tri,
                tri_zone=np.asarray(zone, dtype=np.int32), adj_ptr=adj_ptr,
                adj=adj, is_boundary=is_b, inner=inner, outer=outer, p2io=p2io)
