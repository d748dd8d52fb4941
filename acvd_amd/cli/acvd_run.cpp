// acvd_run: command-line caller of the library with the argument grammar of
// the reference examples (DiscreteRemeshing/Examples/ACVD.cxx:60-131, and
// ACVDQ.cxx:319-338 with -qem 1):
//
//   acvd_run file nvertices gradation [options]
//
//   -q level      ACVD: eigenvalues of the quadric post-placement (default 1,
//                 0 = none, ACVD.cxx:66, 217-291); ACVDQ: quadrics
//                 optimisation level of the metric (default 3)
//   -s threshold  subsampling threshold (default 10); inputs with fewer than
//                 threshold * nvertices vertices are subdivided at their edge
//                 midpoints until they reach it (CheckSubsamplingRatio)
//   -o directory  output directory;  -of file  output file (simplification.ply)
//   -cd file      per-vertex density indicator, one value per line (the
//                 reference reads a VTK image instead); without it and with
//                 gradation > 0, Remesh weights the items with the reference's
//                 curvature indicator (vtkCurvatureMeasure), computed on the
//                 device
//   -qem 0/1      QEM metric with unconstrained initialisation (ACVDQ)
//   -fv file      ACVDQ: fixed vertices, one id per entry (ACVDQ.cxx:246-253);
//                 each becomes a cluster of its own anchored at that vertex
//                 (SetFixedClusters + AnchorItem, :327-334; nvertices counts
//                 the free clusters), and the output is checked to hold every
//                 fixed vertex exactly (:340-366)
//   -ft file      ACVDQ: fixed triangles; their vertices are fixed (:255-281)
//   -w 0/1        energy log: energy.txt in the output directory, one line
//                 "loop seconds energy" per loop and the final energy
//                 (SetWriteToGlobalEnergyLog, ACVD.cxx:157-160)
//   -m 0/1        ForceManifold: restart the minimisation with added clusters
//                 while the output has non-manifold vertices (at most 100 times)
//   -b 0/1        boundary fixing: the output is closed along the input's
//                 boundaries (FixMeshBoundaries)
//   -d -l -cmin -cmax -cf -np -p   accepted; not supported (warned)
//
// Meshes are read from PLY (ascii or binary little-endian) or OBJ and written
// as ASCII PLY.  Polygons are fanned into triangles, as vtkSurface::SetInput
// triangulates its input.  Exit status 0 on success, 1 on error.
#include <cerrno>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "acvd/remeshing.hpp"


namespace {

bool ends_with(const std::string& s, const char* suf) {
  const size_t n = std::strlen(suf);
  return s.size() >= n && s.compare(s.size() - n, n, suf) == 0;
}

void add_polygon(acvd::Mesh& m, const std::vector<int64_t>& idx) {
  for (size_t k = 2; k < idx.size(); ++k) {
    m.triangles.push_back(int32_t(idx[0]));
    m.triangles.push_back(int32_t(idx[k - 1]));
    m.triangles.push_back(int32_t(idx[k]));
  }
}

bool read_obj(const std::string& path, acvd::Mesh& m, std::string& err) {
  std::ifstream in(path);
  if (!in) return err = "cannot open " + path, false;
  std::string line;
  while (std::getline(in, line)) {
    std::istringstream ls(line);
    std::string tag;
    ls >> tag;
    if (tag == "v") {
      double x, y, z;
      if (!(ls >> x >> y >> z)) return err = "bad vertex line: " + line, false;
      m.points.insert(m.points.end(), {x, y, z});
    } else if (tag == "f") {
      std::vector<int64_t> idx;
      std::string tok;
      const int64_t nv = m.GetNumberOfPoints();
      while (ls >> tok) {
        int64_t v = std::strtoll(tok.c_str(), nullptr, 10);  // "a", "a/b", "a//c", "a/b/c"
        v = v < 0 ? nv + v : v - 1;
        if (v < 0 || v >= nv) return err = "face index out of range: " + line, false;
        idx.push_back(v);
      }
      add_polygon(m, idx);
    }
  }
  return true;
}

// Reads one scalar of a PLY type from ascii or little-endian binary input.
bool ply_value(std::istream& in, bool ascii, const std::string& type, double* out) {
  if (ascii) return bool(in >> *out);
  auto rd = [&](auto v) {
    in.read(reinterpret_cast<char*>(&v), sizeof(v));
    *out = double(v);
    return bool(in);
  };
  if (type == "char" || type == "int8") return rd(int8_t());
  if (type == "uchar" || type == "uint8") return rd(uint8_t());
  if (type == "short" || type == "int16") return rd(int16_t());
  if (type == "ushort" || type == "uint16") return rd(uint16_t());
  if (type == "int" || type == "int32") return rd(int32_t());
  if (type == "uint" || type == "uint32") return rd(uint32_t());
  if (type == "float" || type == "float32") return rd(float());
  if (type == "double" || type == "float64") return rd(double());
  return false;
}

bool read_ply(const std::string& path, acvd::Mesh& m, std::string& err) {
  std::ifstream in(path, std::ios::binary);
  if (!in) return err = "cannot open " + path, false;
  struct Prop {
    std::string name, type, count_type;
    bool list = false;
  };
  struct Elem {
    std::string name;
    int64_t count = 0;
    std::vector<Prop> props;
  };
  std::vector<Elem> elems;
  std::string line, format;
  std::getline(in, line);
  if (line.rfind("ply", 0) != 0) return err = path + " is not a PLY file", false;
  while (std::getline(in, line)) {
    if (!line.empty() && line.back() == '\r') line.pop_back();
    std::istringstream ls(line);
    std::string tag;
    ls >> tag;
    if (tag == "format") ls >> format;
    else if (tag == "element") {
      Elem e;
      ls >> e.name >> e.count;
      elems.push_back(e);
    } else if (tag == "property" && !elems.empty()) {
      Prop p;
      std::string t;
      ls >> t;
      if (t == "list") {
        p.list = true;
        ls >> p.count_type >> p.type >> p.name;
      } else {
        p.type = t;
        ls >> p.name;
      }
      elems.back().props.push_back(p);
    } else if (tag == "end_header") break;
  }
  const bool ascii = format == "ascii";
  if (!ascii && format != "binary_little_endian") return err = "unsupported PLY format " + format, false;
  for (const Elem& e : elems) {
    for (int64_t r = 0; r < e.count; ++r) {
      double xyz[3] = {0, 0, 0};
      std::vector<int64_t> idx;
      for (const Prop& p : e.props) {
        if (p.list) {
          double n = 0;
          if (!ply_value(in, ascii, p.count_type, &n)) return err = "truncated PLY list", false;
          idx.clear();
          for (int k = 0; k < int(n); ++k) {
            double v = 0;
            if (!ply_value(in, ascii, p.type, &v)) return err = "truncated PLY list", false;
            idx.push_back(int64_t(v));
          }
        } else {
          double v = 0;
          if (!ply_value(in, ascii, p.type, &v)) return err = "truncated PLY element " + e.name, false;
          if (p.name == "x") xyz[0] = v;
          if (p.name == "y") xyz[1] = v;
          if (p.name == "z") xyz[2] = v;
        }
      }
      if (e.name == "vertex") m.points.insert(m.points.end(), {xyz[0], xyz[1], xyz[2]});
      if (e.name == "face") {
        for (int64_t v : idx)
          if (v < 0 || v >= m.GetNumberOfPoints()) return err = "face index out of range", false;
        add_polygon(m, idx);
      }
    }
  }
  return true;
}

bool write_ply(const std::string& path, const acvd::Mesh& m) {
  FILE* f = std::fopen(path.c_str(), "w");
  if (!f) return false;
  std::fprintf(f, "ply\nformat ascii 1.0\nelement vertex %lld\nproperty double x\nproperty double y\n"
                  "property double z\nelement face %lld\nproperty list uchar int vertex_indices\nend_header\n",
               (long long)m.GetNumberOfPoints(), (long long)m.GetNumberOfCells());
  for (int64_t i = 0; i < m.GetNumberOfPoints(); ++i)
    std::fprintf(f, "%.17g %.17g %.17g\n", m.points[3 * i], m.points[3 * i + 1], m.points[3 * i + 2]);
  for (int64_t t = 0; t < m.GetNumberOfCells(); ++t)
    std::fprintf(f, "3 %d %d %d\n", m.triangles[3 * t], m.triangles[3 * t + 1], m.triangles[3 * t + 2]);
  return std::fclose(f) == 0;
}

// Fixed vertices of -fv (ids) or -ft (triangle ids -> their vertices in
// increasing order), Examples/ACVDQ.cxx:246-281.
bool read_fixed(const std::string& path, bool triangles, const acvd::Mesh& m, std::vector<int32_t>& out,
                std::string& err) {
  std::ifstream in(path);
  if (!in) return err = "cannot open " + path, false;
  out.clear();
  long long id;
  if (!triangles) {
    while (in >> id) {
      if (id < 0 || id >= m.GetNumberOfPoints()) return err = path + ": vertex id out of range", false;
      out.push_back(int32_t(id));
    }
    return true;
  }
  std::vector<char> fixed(size_t(m.GetNumberOfPoints()), 0);
  int n = 0;
  while (in >> id) {
    if (id < 0 || id >= m.GetNumberOfCells()) return err = path + ": triangle id out of range", false;
    for (int k = 0; k < 3; ++k) fixed[size_t(m.triangles[3 * id + k])] = 1;
    ++n;
  }
  for (int64_t i = 0; i < m.GetNumberOfPoints(); ++i)
    if (fixed[size_t(i)]) out.push_back(int32_t(i));
  std::printf("Added %d constraints on triangles\n", n);
  return true;
}

bool read_values(const std::string& path, int64_t n, std::vector<double>& out, std::string& err) {
  std::ifstream in(path);
  if (!in) return err = "cannot open " + path, false;
  out.clear();
  double v;
  while (in >> v) out.push_back(v);
  if (int64_t(out.size()) != n) return err = path + ": expected one value per vertex", false;
  return true;
}

template <class Remesher>
int run(Remesher* remesh, const acvd::Mesh& mesh, int K, double gradation, int subsampling, int q_level,
        bool qem, const std::vector<double>* indicator, const std::vector<int32_t>* fixed,
        const std::string& out_path, const std::string& smooth_path, int energy_log, const std::string& prefix,
        int force_manifold, int boundary_fixing) {
  remesh->SetInput(&mesh);
  if (boundary_fixing) {  // ACVD.cxx:187-190, ACVDQ.cxx:241-245
    std::printf("Setting boundary fixing to %d\n", boundary_fixing);
    remesh->SetBoundaryFixing(boundary_fixing);
  }
  if (force_manifold) {  // ACVD.cxx:118-120, ACVDQ.cxx:159-162
    std::printf("Force Manifold=%d\n", force_manifold);
    remesh->SetForceManifold(true);
  }
  if (energy_log) {  // ACVD.cxx:157-160, ACVDQ.cxx:199-203; energy.txt beside the outputs
    std::printf("Setting writing energy log file to %d\n", energy_log);
    remesh->SetWriteToGlobalEnergyLog(energy_log);
    remesh->SetOutputDirectory(prefix.c_str());
  }
  if (fixed) {
    remesh->SetFixedClusters(fixed);
    remesh->SetNumberOfClusters(K + int(fixed->size()));
    std::printf("Read %zu fixed Ids\n", fixed->size());
  } else {
    remesh->SetNumberOfClusters(K);
  }
  remesh->SetConsoleOutput(2);
  remesh->SetSubsamplingThreshold(subsampling);
  remesh->GetMetric()->SetGradation(gradation);
  if (indicator) remesh->GetMetric()->SetCurvatureInfo(indicator);
  if (qem) {
    remesh->GetMetric()->SetQuadricsOptimizationLevel(q_level);
    remesh->SetUnconstrainedInitialization(1);
  }
  if (!remesh->Remesh()) {
    std::fprintf(stderr, "acvd_run: %s\n", remesh->GetLastError().c_str());
    return 1;
  }
  if (!qem && q_level != 0) {
    // ACVD.cxx:217-291: the clustering's mesh is written as smooth_<file>,
    // then the vertices take the level-q quadric placement
    if (!write_ply(smooth_path, *remesh->GetOutput())) {
      std::fprintf(stderr, "acvd_run: cannot write %s\n", smooth_path.c_str());
      return 1;
    }
    if (!remesh->BuildOutput(q_level)) {
      std::fprintf(stderr, "acvd_run: %s\n", remesh->GetLastError().c_str());
      return 1;
    }
  }
  const acvd::Mesh* out = remesh->GetOutput();
  if (fixed) {  // the vertex constraints are respected (ACVDQ.cxx:340-366)
    for (size_t i = 0; i < fixed->size(); ++i) {
      double c1[3], c2[3];
      mesh.GetPoint((*fixed)[i], c1);
      out->GetPoint(int64_t(i), c2);
      for (int j = 0; j < 3; ++j)
        if (c1[j] != c2[j]) {
          std::fprintf(stderr, "Error, vertex %d has been lost\n", (*fixed)[i]);
          return 1;
        }
    }
    std::printf("Constraints on vertices have been checked\n");
  }
  if (!write_ply(out_path, *out)) {
    std::fprintf(stderr, "acvd_run: cannot write %s\n", out_path.c_str());
    return 1;
  }
  std::printf("output: %lld vertices, %lld triangles -> %s\n", (long long)out->GetNumberOfPoints(),
              (long long)out->GetNumberOfCells(), out_path.c_str());
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc < 4) {
    std::printf("Usage : acvd_run file nvertices gradation [options]\n"
                "nvertices is the desired number of vertices\n"
                "gradation defines the influence of local curvature (0=uniform meshing)\n"
                "options: -q level, -s threshold, -o directory, -of file, -cd indicator file, -qem 0/1,\n"
                "         -fv fixed vertices file, -ft fixed triangles file (with -qem 1), -w energy log 0/1,\n"
                "         -m force manifold 0/1, -b boundary fixing 0/1\n");
    return 0;
  }
  const std::string input = argv[1];
  const int K = std::atoi(argv[2]);
  const double gradation = std::atof(argv[3]);
  int subsampling = 10, q_level = -1, qem = 0, energy_log = 0, force_manifold = 0, boundary_fixing = 0;
  std::string out_dir, out_file = "simplification.ply", density, fixed_file;
  bool fixed_triangles = false;
  for (int a = 4; a + 1 < argc; a += 2) {
    const std::string key = argv[a], value = argv[a + 1];
    if (key == "-s") subsampling = std::atoi(value.c_str());
    else if (key == "-q") q_level = std::atoi(value.c_str());
    else if (key == "-o") out_dir = value;
    else if (key == "-of") out_file = value;
    else if (key == "-cd") density = value;
    else if (key == "-qem") qem = std::atoi(value.c_str());
    else if (key == "-w") energy_log = std::atoi(value.c_str());
    else if (key == "-m") force_manifold = std::atoi(value.c_str());
    else if (key == "-b") boundary_fixing = std::atoi(value.c_str());
    else if (key == "-fv" || key == "-ft") {
      fixed_file = value;
      fixed_triangles = key == "-ft";
    }
    else std::printf("option %s %s is not supported here; ignored\n", key.c_str(), value.c_str());
  }
  if (q_level < 0) q_level = qem ? 3 : 1;  // ACVDQ / ACVD defaults
  acvd::Mesh mesh;
  std::string err;
  const bool read_ok = ends_with(input, ".obj") ? read_obj(input, mesh, err) : read_ply(input, mesh, err);
  if (!read_ok) {
    std::fprintf(stderr, "acvd_run: %s\n", err.c_str());
    return 1;
  }
  std::printf("load : %s (%lld vertices, %lld triangles)\n", input.c_str(), (long long)mesh.GetNumberOfPoints(),
              (long long)mesh.GetNumberOfCells());
  std::vector<double> indicator;
  const std::vector<double>* ind = nullptr;
  if (!density.empty()) {
    if (!read_values(density, mesh.GetNumberOfPoints(), indicator, err)) {
      std::fprintf(stderr, "acvd_run: %s\n", err.c_str());
      return 1;
    }
    ind = &indicator;
  }
  std::vector<int32_t> fixed;
  if (!fixed_file.empty()) {
    if (!qem) {
      std::fprintf(stderr, "acvd_run: -fv / -ft need the QEM metric (-qem 1), as in ACVDQ\n");
      return 1;
    }
    if (!read_fixed(fixed_file, fixed_triangles, mesh, fixed, err)) {
      std::fprintf(stderr, "acvd_run: %s\n", err.c_str());
      return 1;
    }
  }
  const std::string prefix = out_dir.empty() ? "" : out_dir + "/";
  const std::string out_path = prefix + out_file, smooth_path = prefix + "smooth_" + out_file;
  int rc;
  if (qem) {
    auto* r = acvd::QIsotropicDiscreteRemeshing::New();
    rc = run(r, mesh, K, gradation, subsampling, q_level, true, ind, fixed_file.empty() ? nullptr : &fixed, out_path,
             smooth_path, energy_log, prefix, force_manifold, boundary_fixing);
    r->Delete();
  } else {
    auto* r = acvd::IsotropicDiscreteRemeshing::New();
    rc = run(r, mesh, K, gradation, subsampling, q_level, false, ind, nullptr, out_path, smooth_path, energy_log,
             prefix, force_manifold, boundary_fixing);
    r->Delete();
  }
  return rc;
}
