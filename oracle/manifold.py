"""TEST INFRASTRUCTURE — CPU oracle, not part of the product.

Pure-Python restatement of the topology tests behind Remesh's ForceManifold
loop, written after the reference's own data structure (small meshes only):

  vtkSurfaceBase edge table        Common/vtkSurfaceBase.cxx:1166-1221 (AddFace:
                                   edges (v0,v1), (v1,v2), (v2,v0) of each face,
                                   an edge created at its first sighting and
                                   appended to both vertex edge lists; the
                                   faces of an edge fill Poly1, Poly2, then the
                                   non-manifold list)
  IsEdgeManifold                   Common/vtkSurfaceBase.h:521-528
  IsVertexManifold                 Common/vtkSurfaceBase.cxx:259-317 (the walk
                                   around the vertex, Conquer
                                   Common/vtkSurfaceBase.h:402-425,
                                   GetThirdPoint :383-400)
  BuildDelaunayTriangulation's     DiscreteRemeshing/vtkDiscreteRemeshing.h:
  wire edges under ForceManifold   1114-1131
  DetectNonManifoldOutputVertices  DiscreteRemeshing/vtkDiscreteRemeshing.h:165-383
  vtkIdList::DeleteId              every occurrence replaced by the last id
"""


class Surface:
    def __init__(self, n_vertices, faces):
        self.faces = [tuple(int(x) for x in f) for f in faces]
        self.vertex_edges = [[] for _ in range(n_vertices)]
        self.edge_vertices = []
        self.edge_faces = []  # [Poly1, Poly2, [more...]]
        self.edge_of = {}
        for fid, f in enumerate(self.faces):
            for j in range(3):
                a, b = f[j], f[(j + 1) % 3]
                e = self.add_edge(a, b)
                ef = self.edge_faces[e]
                if ef[0] < 0:
                    ef[0] = fid
                elif ef[1] < 0:
                    ef[1] = fid
                else:
                    ef[2].append(fid)

    def add_edge(self, a, b):
        k = (min(a, b), max(a, b))
        e = self.edge_of.get(k)
        if e is None:
            e = len(self.edge_vertices)
            self.edge_of[k] = e
            self.edge_vertices.append((a, b))
            self.edge_faces.append([-1, -1, []])
            self.vertex_edges[a].append(e)
            self.vertex_edges[b].append(e)
        return e

    def is_edge(self, a, b):
        return self.edge_of.get((min(a, b), max(a, b)), -1)

    def edge_manifold(self, e):
        ef = self.edge_faces[e]
        return not (ef[1] < 0 or ef[2])

    def third_point(self, f, v1, v2):
        v3, v4, v5 = self.faces[f]
        if v1 != v3 and v2 != v3:
            return v3
        if v1 != v4 and v2 != v4:
            return v4
        return v5

    def conquer(self, f1, v1, v2):
        e = self.is_edge(v1, v2)
        if e < 0:
            return -1, -1
        f2 = self.edge_faces[e][1]
        if f2 == -1:
            return -1, -1
        if f2 == f1:
            f2 = self.edge_faces[e][0]
        return f2, self.third_point(f2, v1, v2)

    def neighbours(self, v):
        out = []
        for e in self.vertex_edges[v]:
            a, b = self.edge_vertices[e]
            out.append(b if a == v else a)
        return out

    def is_vertex_manifold(self, v):
        edges = self.vertex_edges[v]
        remaining = len(edges)
        if remaining < 2:
            return False
        for e in edges:
            if not self.edge_manifold(e):
                return False
        first = edges[0]
        a, b = self.edge_vertices[first]
        first_vertex = b if a == v else a
        remaining -= 1
        f1, f2 = self.edge_faces[first][0], self.edge_faces[first][1]
        v2 = self.third_point(f1, v, first_vertex)
        while True:
            remaining -= 1
            if remaining == 0:
                return True
            f3, v1 = self.conquer(f1, v, v2)
            f1, v2 = f3, v1
            if not (f1 >= 0 and v2 != first_vertex):
                break
        if f2 < 0 or v2 == first_vertex:
            return False
        v2 = self.third_point(f2, v, first_vertex)
        while True:
            remaining -= 1
            if remaining == 0:
                return True
            f3, v1 = self.conquer(f2, v, v2)
            f2, v2 = f3, v1
            if not (f2 >= 0 and v2 != first_vertex):
                break
        return False


def delete_id(lst, x):
    i = 0
    while i < len(lst):
        while i < len(lst) and lst[i] != x:
            i += 1
        if i < len(lst):
            lst[i] = lst[-1]
            lst.pop()


def force_manifold_step(n_vertices, tri, labels, K, out_tri):
    """DetectNonManifoldOutputVertices after a ForceManifold output build:
    returns (issues, new labels, new K, frozen flags)."""
    inp = Surface(n_vertices, tri)
    out = Surface(K, out_tri)
    lab = [int(x) for x in labels]
    # wire edges between clusters adjacent in the input (input edges in order)
    for (i1, i2) in inp.edge_vertices:
        c1, c2 = lab[i1], lab[i2]
        if c1 != c2 and 0 <= c1 < K and 0 <= c2 < K and out.is_edge(c1, c2) < 0:
            out.add_edge(c1, c2)
    items = [None] * K
    for i, c in enumerate(lab):
        if 0 <= c < K:
            if items[c] is None:
                items[c] = []
            items[c].append(i)
    frozen = [1] * K
    issues = []
    for c in range(K):
        if out.is_vertex_manifold(c):
            continue
        if items[c] is None:
            continue
        if any(not inp.is_vertex_manifold(i) for i in items[c]):
            continue
        issues.append(c)
        frozen[c] = 0
        for n in out.neighbours(c):
            frozen[n] = 0
    Kn = K
    for c in issues:
        lst = items[c]
        nc = Kn
        items.append(None)
        frozen.append(0)
        Kn += 1
        if len(lst) > 1:
            it = lst[0]
            lab[it] = nc
            items[nc] = [it]
            delete_id(lst, it)
            continue
        it = lst[0]
        found = False
        for u in inp.neighbours(it):
            uc = lab[u]
            if 0 <= uc < len(items) and items[uc] is not None and len(items[uc]) > 1:
                lab[u] = nc
                delete_id(items[uc], u)
                items[nc] = [u]
                found = True
                break
        if not found:
            items.pop()
            frozen.pop()
            Kn -= 1
    return len(issues), lab, Kn, frozen


def fix_mesh_boundaries(n_vertices, xyz, tri, labels, K):
    """FixMeshBoundaries (vtkDiscreteRemeshing.h:425-579) with BoundaryFixing
    on: (added vertices, added triangles); added vertex j has output id K + j.
    GetNumberOfBoundaries as the reference computes it
    (Common/vtkSurfaceBase.cxx:939-949: Edges[i] for i < valence are the
    mesh's first edges, not the vertex's)."""
    s = Surface(n_vertices, tri)
    ne = len(s.edge_vertices)
    lab = [int(x) for x in labels]
    newp = [-2] * ne
    add_xyz, add_tri = [], []

    def boundary(e):
        return s.edge_faces[e][1] < 0 and s.edge_faces[e][0] >= 0

    for e in range(ne):
        if not boundary(e):
            continue
        v1, v2 = s.edge_vertices[e]
        c1, c2 = lab[v1], lab[v2]
        if c1 != c2 and 0 <= c1 < K and 0 <= c2 < K:
            newp[e] = K + len(add_xyz)
            add_xyz.append([0.5 * (xyz[v1][k] + xyz[v2][k]) for k in range(3)])
            add_tri.append((newp[e], c1, c2))
    for i in range(ne):
        if newp[i] >= 0 or not boundary(i):
            continue
        cluster = lab[s.edge_vertices[i][0]]
        queue = [i]
        elist = []
        while queue:
            e = queue.pop(0)
            st = newp[e]
            if st == -2:
                newp[e] = -1
                for v in s.edge_vertices[e]:
                    for f in s.vertex_edges[v]:
                        if s.edge_faces[f][1] < 0:
                            queue.append(f)
            elif st >= 0 and e not in elist:
                elist.append(e)
        if len(elist) >= 2 and 0 <= cluster < K:
            add_tri.append((cluster, newp[elist[0]], newp[elist[1]]))
    for v in range(n_vertices):
        deg = len(s.vertex_edges[v])
        nb = sum(1 for k in range(min(deg, ne)) if s.edge_faces[k][1] < 0)
        if nb // 2 <= 0:
            continue
        elist = []
        for e in s.vertex_edges[v]:
            if newp[e] >= 0 and newp[e] not in elist:
                elist.append(newp[e])
        if len(elist) > 1 and 0 <= lab[v] < K:
            add_tri.append((lab[v], elist[0], elist[1]))
    return add_xyz, add_tri
