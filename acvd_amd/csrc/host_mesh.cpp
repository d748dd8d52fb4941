// Host-side mesh preparation of Remesh() (vtkDiscreteRemeshing::
// CheckSubsamplingRatio, DiscreteRemeshing/vtkDiscreteRemeshing.h:840-875):
// one level of midpoint subdivision (vtkSurface::Subdivide,
// Common/vtkSurface.cxx:605-677).  Runs once per Remesh on meshes too coarse
// for the requested cluster count, before the device build.
#include "host_mesh.h"

#include <algorithm>
#include <cstdio>
#include <vector>

#include "host_parallel.h"

namespace acvd {

void edge_table(int64_t nv, int64_t nf, const int32_t* tri, EdgeTable* out) {
  EdgeTable& E = *out;
  // edges in creation order: faces in order (inactive faces, whose first two
  // vertices coincide, create none), edges (v0,v1), (v1,v2), (v2,v0), a new
  // edge at its first sighting (Common/vtkSurfaceBase.cxx:1166-1221, 1443-1451).
  // Each edge is looked up in the list of its smaller end.
  std::vector<int64_t> cap(nv + 1, 0);
  for (int64_t f = 0; f < nf; ++f) {
    const int32_t* t = tri + 3 * f;
    if (t[0] == t[1]) continue;
    for (int j = 0; j < 3; ++j) {
      const int32_t a = t[j], b = t[(j + 1) % 3];
      if (a != b) cap[std::min(a, b) + 1]++;
    }
  }
  for (int64_t v = 0; v < nv; ++v) cap[v + 1] += cap[v];
  std::vector<int32_t> other(size_t(cap[nv])), eid(size_t(cap[nv]));
  std::vector<int64_t> used(nv, 0);
  E.face_edge.assign(size_t(3 * nf), -1);
  E.ev1.clear();
  E.ev2.clear();
  E.poly1.clear();
  E.poly2.clear();
  for (int64_t f = 0; f < nf; ++f) {
    const int32_t* t = tri + 3 * f;
    if (t[0] == t[1]) continue;
    for (int j = 0; j < 3; ++j) {
      const int32_t a = t[j], b = t[(j + 1) % 3];
      if (a == b) continue;
      const int32_t lo = std::min(a, b), hi = std::max(a, b);
      const int64_t s0 = cap[lo];
      int32_t id = -1;
      for (int64_t k = s0; k < s0 + used[lo]; ++k)
        if (other[size_t(k)] == hi) {
          id = eid[size_t(k)];
          break;
        }
      if (id < 0) {
        id = int32_t(E.ev1.size());
        E.ev1.push_back(a);
        E.ev2.push_back(b);
        E.poly1.push_back(-1);
        E.poly2.push_back(-1);
        other[size_t(s0 + used[lo])] = hi;
        eid[size_t(s0 + used[lo])] = id;
        used[lo]++;
      }
      E.face_edge[size_t(3 * f + j)] = id;
      // the faces of an edge: Poly1, Poly2, then the non-manifold ones
      if (E.poly1[size_t(id)] < 0) E.poly1[size_t(id)] = int32_t(f);
      else if (E.poly2[size_t(id)] < 0) E.poly2[size_t(id)] = int32_t(f);
    }
  }
  // edges of each vertex in creation order (appended to both ends)
  const int64_t ne = int64_t(E.ev1.size());
  E.vptr.assign(size_t(nv + 1), 0);
  for (int64_t e = 0; e < ne; ++e) {
    E.vptr[size_t(E.ev1[size_t(e)] + 1)]++;
    E.vptr[size_t(E.ev2[size_t(e)] + 1)]++;
  }
  for (int64_t v = 0; v < nv; ++v) E.vptr[size_t(v + 1)] += E.vptr[size_t(v)];
  E.vedge.assign(size_t(E.vptr[size_t(nv)]), 0);
  std::vector<int64_t> fill(E.vptr.begin(), E.vptr.end() - 1);
  for (int64_t e = 0; e < ne; ++e) {
    E.vedge[size_t(fill[size_t(E.ev1[size_t(e)])]++)] = int32_t(e);
    E.vedge[size_t(fill[size_t(E.ev2[size_t(e)])]++)] = int32_t(e);
  }
}

int64_t subdivide_mesh(int64_t nv, const double* xyz, int64_t nf, const int32_t* tri, std::vector<double>* out_xyz,
                       std::vector<int32_t>* out_tri, std::vector<int32_t>* parents) {
  EdgeTable E;
  edge_table(nv, nf, tri, &E);
  const std::vector<int32_t>& ev1 = E.ev1;
  const std::vector<int32_t>& ev2 = E.ev2;
  const std::vector<int32_t>& face_edge = E.face_edge;
  const int64_t ne = int64_t(ev1.size());
  // the input vertices, then one vertex per edge at its midpoint
  out_xyz->assign(xyz, xyz + 3 * nv);
  out_xyz->resize(size_t(3 * (nv + ne)));
  parents->resize(size_t(2 * ne));
  for (int64_t e = 0; e < ne; ++e) {
    const double* p1 = xyz + 3 * int64_t(ev1[size_t(e)]);
    const double* p2 = xyz + 3 * int64_t(ev2[size_t(e)]);
    for (int k = 0; k < 3; ++k) (*out_xyz)[size_t(3 * (nv + e) + k)] = 0.5 * (p1[k] + p2[k]);
    (*parents)[size_t(2 * e)] = ev1[size_t(e)];
    (*parents)[size_t(2 * e + 1)] = ev2[size_t(e)];
  }
  // four faces per active face: (V1,V4,V6), (V4,V2,V5), (V5,V3,V6), (V4,V5,V6)
  // with V4, V5, V6 the children of edges (V1,V2), (V2,V3), (V3,V1)
  out_tri->clear();
  out_tri->reserve(size_t(12 * nf));
  for (int64_t f = 0; f < nf; ++f) {
    const int32_t* t = tri + 3 * f;
    if (t[0] == t[1]) continue;
    const int32_t v1 = t[0], v2 = t[1], v3 = t[2];
    const int32_t v4 = int32_t(nv + face_edge[size_t(3 * f)]);
    const int32_t v5 = int32_t(nv + face_edge[size_t(3 * f + 1)]);
    const int32_t v6 = int32_t(nv + face_edge[size_t(3 * f + 2)]);
    const int32_t nt[12] = {v1, v4, v6, v4, v2, v5, v5, v3, v6, v4, v5, v6};
    out_tri->insert(out_tri->end(), nt, nt + 12);
  }
  return ne;
}

namespace {

// Faces of each vertex and, per vertex, the manifold test over them.
bool manifold_at(int64_t v, const int64_t* vfptr, const int32_t* vf, const int32_t* tri, std::vector<int32_t>& nb,
                 std::vector<int32_t>& cnt, std::vector<int32_t>& parent) {
  nb.clear();
  cnt.clear();
  auto index_of = [&](int32_t u) {
    for (size_t q = 0; q < nb.size(); ++q)
      if (nb[q] == u) return int(q);
    nb.push_back(u);
    cnt.push_back(0);
    return int(nb.size() - 1);
  };
  // edges at v and the faces through each (a face with v once adds one to
  // each of its two edges at v)
  for (int64_t k = vfptr[v]; k < vfptr[v + 1]; ++k) {
    const int32_t* t = tri + 3 * int64_t(vf[k]);
    for (int j = 0; j < 3; ++j) {
      if (t[j] != v) continue;
      const int32_t a = t[(j + 1) % 3], b = t[(j + 2) % 3];
      if (a != v) cnt[size_t(index_of(a))]++;
      if (b != v && b != a) cnt[size_t(index_of(b))]++;
    }
  }
  if (nb.size() < 2) return false;
  for (int32_t c : cnt)
    if (c != 2) return false;
  // one fan: the faces join the ring neighbours into one component
  parent.resize(nb.size());
  for (size_t q = 0; q < nb.size(); ++q) parent[q] = int32_t(q);
  auto find = [&](int32_t x) {
    while (parent[size_t(x)] != x) x = parent[size_t(x)] = parent[size_t(parent[size_t(x)])];
    return x;
  };
  for (int64_t k = vfptr[v]; k < vfptr[v + 1]; ++k) {
    const int32_t* t = tri + 3 * int64_t(vf[k]);
    for (int j = 0; j < 3; ++j) {
      if (t[j] != v) continue;
      const int32_t a = find(index_of(t[(j + 1) % 3])), b = find(index_of(t[(j + 2) % 3]));
      if (a != b) parent[size_t(a)] = b;
    }
  }
  const int32_t r = find(0);
  for (size_t q = 1; q < nb.size(); ++q)
    if (find(int32_t(q)) != r) return false;
  return true;
}

// vtkIdList::DeleteId: every occurrence is replaced by the list's last id.
void delete_id(std::vector<int32_t>& list, int32_t id) {
  size_t i = 0;
  while (i < list.size()) {
    for (; i < list.size(); ++i)
      if (list[i] == id) break;
    if (i < list.size()) {
      list[i] = list.back();
      list.pop_back();
    }
  }
}

}  // namespace

void vertex_manifold(int64_t nv, int64_t nf, const int32_t* tri, std::vector<uint8_t>* ok) {
  VertexTopology topo;
  vertex_topology(nv, nf, tri, &topo);
  ok->assign(size_t(nv), 0);
  parallel_for(nv, [&](int64_t v0, int64_t v1) {
    std::vector<int32_t> nb, cnt, parent;
    for (int64_t v = v0; v < v1; ++v)
      (*ok)[size_t(v)] = manifold_at(v, topo.vfptr.data(), topo.vf.data(), tri, nb, cnt, parent) ? 1 : 0;
  });
}

int64_t force_manifold_step(int64_t nv, const VertexTopology& topo, const std::vector<uint8_t>& in_ok,
                            int64_t n_out, const int32_t* out_tri, int32_t* K_io, int32_t* labels,
                            std::vector<int32_t>* frozen, bool verbose) {
  const int32_t K = *K_io;
  // output edges: those of the triangles, then the wire edges between
  // clusters adjacent in the input that no triangle joins
  std::vector<uint64_t> tri_edges;
  tri_edges.reserve(size_t(3 * n_out));
  auto key = [](int32_t a, int32_t b) {
    return (uint64_t(uint32_t(std::min(a, b))) << 32) | uint64_t(uint32_t(std::max(a, b)));
  };
  for (int64_t f = 0; f < n_out; ++f)
    for (int j = 0; j < 3; ++j) tri_edges.push_back(key(out_tri[3 * f + j], out_tri[3 * f + (j + 1) % 3]));
  std::sort(tri_edges.begin(), tri_edges.end());
  tri_edges.erase(std::unique(tri_edges.begin(), tri_edges.end()), tri_edges.end());
  std::vector<uint8_t> wire(size_t(K), 0);
  std::vector<uint64_t> wire_edges;
  for (int64_t v = 0; v < nv; ++v)
    for (int64_t r = topo.rptr[size_t(v)]; r < topo.rptr[size_t(v + 1)]; ++r) {
      const int32_t u = topo.ring[size_t(r)];
      if (u <= v) continue;
      const int32_t c1 = labels[v], c2 = labels[u];
      if (c1 == c2 || c1 < 0 || c1 >= K || c2 < 0 || c2 >= K) continue;
      const uint64_t e = key(c1, c2);
      if (std::binary_search(tri_edges.begin(), tri_edges.end(), e)) continue;
      wire[size_t(c1)] = wire[size_t(c2)] = 1;
      wire_edges.push_back(e);
    }
  // output vertex neighbours (GetVertexNeighbours of the output: both kinds)
  std::vector<uint64_t> all(tri_edges);
  all.insert(all.end(), wire_edges.begin(), wire_edges.end());
  std::sort(all.begin(), all.end());
  all.erase(std::unique(all.begin(), all.end()), all.end());
  std::vector<int64_t> nptr(size_t(K) + 1, 0);
  for (uint64_t e : all) {
    nptr[size_t(e >> 32) + 1]++;
    nptr[size_t(e & 0xffffffffu) + 1]++;
  }
  for (int32_t c = 0; c < K; ++c) nptr[size_t(c) + 1] += nptr[size_t(c)];
  std::vector<int32_t> nbr(size_t(nptr[size_t(K)]));
  {
    std::vector<int64_t> fill(nptr.begin(), nptr.end() - 1);
    for (uint64_t e : all) {
      const int32_t a = int32_t(e >> 32), b = int32_t(e & 0xffffffffu);
      nbr[size_t(fill[size_t(a)]++)] = b;
      nbr[size_t(fill[size_t(b)]++)] = a;
    }
  }
  // manifold test of the output vertices over the triangles; a wire edge
  // (no face) makes both its ends non-manifold
  std::vector<uint8_t> out_ok;
  vertex_manifold(K, n_out, out_tri, &out_ok);
  // items of each cluster, in increasing item order
  std::vector<std::vector<int32_t>> items(static_cast<size_t>(K));
  int64_t misassigned = 0;
  for (int64_t i = 0; i < nv; ++i) {
    const int32_t c = labels[i];
    if (c < 0 || c >= K) misassigned++;
    else items[size_t(c)].push_back(int32_t(i));
  }
  if (verbose) std::printf("Starting detection of non-manifold vertices\n");
  if (misassigned && verbose) std::printf("%lld uncorrectly associated items\n", (long long)misassigned);
  frozen->assign(size_t(K), 1);
  std::vector<int32_t> issues;
  for (int32_t c = 0; c < K; ++c) {
    if (out_ok[size_t(c)] && !wire[size_t(c)]) continue;
    if (verbose) std::printf("Cluster %d is non manifold\n", c);
    if (items[size_t(c)].empty()) {
      if (verbose) std::printf(".... but empty. Skipping\n");
      continue;
    }
    bool problem = true;
    for (int32_t it : items[size_t(c)])
      if (!in_ok[size_t(it)]) {
        problem = false;
        if (verbose) std::printf("discarding this topology issue as the input mesh also has a topology issue here\n");
        break;
      }
    if (!problem) continue;
    issues.push_back(c);
    (*frozen)[size_t(c)] = 0;
    for (int64_t q = nptr[size_t(c)]; q < nptr[size_t(c) + 1]; ++q) (*frozen)[size_t(nbr[size_t(q)])] = 0;
  }
  int32_t Kn = K;
  for (int32_t c : issues) {
    std::vector<int32_t>& list = items[size_t(c)];
    const int32_t nc = Kn;
    if (list.size() > 1) {
      const int32_t it = list[0];
      labels[it] = nc;
      delete_id(list, it);
      items.push_back({it});
      frozen->push_back(0);
      ++Kn;
      continue;
    }
    // one item: take a ring neighbour from a cluster of more than one item
    const int32_t it = list[0];
    bool found = false;
    for (int64_t r = topo.rptr[size_t(it)]; r < topo.rptr[size_t(it) + 1]; ++r) {
      const int32_t u = topo.ring[size_t(r)];
      const int32_t uc = labels[u];
      if (uc < 0 || uc >= Kn) continue;
      if (items[size_t(uc)].size() > 1) {
        labels[u] = nc;
        delete_id(items[size_t(uc)], u);
        items.push_back({u});
        frozen->push_back(0);
        ++Kn;
        found = true;
        break;
      }
    }
    if (!found && verbose) std::printf("Could not find a place to add cluster %d near cluster %d\n", nc, c);
  }
  *K_io = Kn;
  return int64_t(issues.size());
}

}  // namespace acvd

namespace acvd {

bool fix_mesh_boundaries(int64_t nv, const double* xyz, int64_t nf, const int32_t* tri, const int32_t* labels,
                         int32_t K, std::vector<double>* out_xyz, std::vector<int32_t>* out_tri) {
  EdgeTable E;
  edge_table(nv, nf, tri, &E);
  const int64_t ne = int64_t(E.ev1.size());
  auto cluster = [&](int32_t v) { return labels[v]; };
  auto valid = [&](int32_t c) { return c >= 0 && c < K; };
  auto boundary = [&](int64_t e) { return E.poly2[size_t(e)] < 0 && E.poly1[size_t(e)] >= 0; };
  // -2 never visited, -1 visited, >= 0 the output vertex at the edge's midpoint
  std::vector<int32_t> newp(size_t(ne), -2);
  bool needed = false;
  for (int64_t e = 0; e < ne; ++e) {
    if (!boundary(e)) continue;
    const int32_t v1 = E.ev1[size_t(e)], v2 = E.ev2[size_t(e)];
    const int32_t c1 = cluster(v1), c2 = cluster(v2);
    if (c1 == c2 || !valid(c1) || !valid(c2)) continue;
    const int32_t p = int32_t(out_xyz->size() / 3);
    for (int k = 0; k < 3; ++k) out_xyz->push_back(0.5 * (xyz[3 * int64_t(v1) + k] + xyz[3 * int64_t(v2) + k]));
    newp[size_t(e)] = p;
    needed = true;
    out_tri->insert(out_tri->end(), {p, c1, c2});
  }
  // runs of boundary edges inside one cluster, closed by the two cut edges
  // met at their ends (the walk stops at cut edges and collects them)
  std::vector<int32_t> elist;
  std::vector<int32_t> queue;
  auto insert_unique = [&](int32_t x) {
    if (std::find(elist.begin(), elist.end(), x) == elist.end()) elist.push_back(x);
  };
  for (int64_t i = 0; i < ne; ++i) {
    if (newp[size_t(i)] >= 0 || newp[size_t(i)] == -1 || !boundary(i)) continue;
    const int32_t c = cluster(E.ev1[size_t(i)]);
    queue.assign(1, int32_t(i));
    size_t head = 0;
    elist.clear();
    while (head < queue.size()) {
      const int32_t e = queue[head++];
      const int32_t st = newp[size_t(e)];
      if (st == -2) {
        newp[size_t(e)] = -1;
        for (const int32_t v : {E.ev1[size_t(e)], E.ev2[size_t(e)]})
          for (int64_t q = E.vptr[size_t(v)]; q < E.vptr[size_t(v) + 1]; ++q) {
            const int32_t f = E.vedge[size_t(q)];
            if (E.poly2[size_t(f)] < 0) queue.push_back(f);
          }
      } else if (st >= 0) {
        insert_unique(int32_t(e));
      }
    }
    if (elist.size() >= 2 && valid(c))
      out_tri->insert(out_tri->end(), {c, newp[size_t(elist[0])], newp[size_t(elist[1])]});
  }
  // a vertex between two cut edges (GetNumberOfBoundaries as the reference
  // computes it, Common/vtkSurfaceBase.cxx:939-949: it tests the global
  // edges 0 .. valence - 1, not the vertex's own edges)
  for (int64_t v = 0; v < nv; ++v) {
    const int64_t deg = E.vptr[size_t(v) + 1] - E.vptr[size_t(v)];
    int64_t nb = 0;
    for (int64_t k = 0; k < deg && k < ne; ++k) nb += E.poly2[size_t(k)] < 0;
    if (nb / 2 <= 0) continue;
    elist.clear();
    for (int64_t q = E.vptr[size_t(v)]; q < E.vptr[size_t(v) + 1]; ++q) {
      const int32_t p = newp[size_t(E.vedge[size_t(q)])];
      if (p >= 0) insert_unique(p);
    }
    if (elist.size() > 1 && valid(cluster(int32_t(v))))
      out_tri->insert(out_tri->end(), {cluster(int32_t(v)), elist[0], elist[1]});
  }
  return needed;
}

}  // namespace acvd
