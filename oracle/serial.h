// TEST INFRASTRUCTURE — CPU oracle, not part of the product.
//
// Oracle A: serial restatement of the reference's clustering engine on the
// vertex path, i.e. vtkVerticesProcessing< vtkDiscreteRemeshing< Metric > >
// driven through vtkUniformClustering (Common/vtkUniformClustering.h):
//   SetNumberOfClusters      :62-69
//   ReComputeClustersSize    :352-373
//   ReComputeStatistics      :375-403
//   CleanClustering          :405-549
//   FillHolesInClustering    :551-633
//   FillQueuesFromClustering :635-652
//   ProcessClustering        :654-714
//   SetAllClustersToModified :716-722
//   MinimizeEnergy           :724-830
//   ProcessOneLoop           :832-995
//   InitSamples              :1123-1175 (type 1 and 2; type 0 uses VTK's RNG,
//                                        restated with our own stream)
//   ComputeInitialRandomSampling :1177-1316
//   ComputeGlobalEnergy      :1318-1346 (global_energy), WriteToGlobalEnergyLog :1383-1395
//   Init / Allocate          :1348-1381, defaults :1397-1419
// and the vertex adaptor DiscreteRemeshing/vtkVerticesProcessing.h:
//   AddItemRingToProcess :142-148, GetEdgeItems :150-153,
//   ConnexityConstraintProblemLocal :168-237.
#pragma once

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdint>
#include <deque>
#include <queue>
#include <random>
#include <vector>

#include "metrics.h"
#include "ref_mesh.h"

namespace oracle {

// Optional trace of a run (tests): one record per evaluated cut edge (the
// three energies of :922-956 and the decision of :960-991: 0 stay, 1 move I1,
// 2 move I2), one per loop (the state the loop ran with and what followed it).
struct EdgeTrace {
  int32_t loop, decision;
  int64_t edge, i1, i2;
  double try1, try2, try3;
};
struct LoopTrace {
  int32_t loop, mods, connexity, constrained, convergence, disconnected;
};

// vtkMath's uniform sequence (vtkMinimalStandardRandomSequence): Park and
// Miller's minimal standard x' = 16807 x mod (2^31 - 1), computed with
// Schrage's method; SetSeed clamps the seed into [1, 2^31 - 2] and steps
// three times; each draw steps, then reads x / (2^31 - 1).
struct VtkMinStd {
  int64_t x = 1;
  void step() { x = (16807 * x) % 2147483647; }
  void set_seed(int32_t s) {
    x = s;
    if (x < 1) x += 2147483646;
    else if (x == 2147483647) x = 1;
    step();
    step();
    step();
  }
  double uniform() {
    step();
    return double(x) / 2147483647.0;
  }
  double range(double a, double b) {
    step();
    return a == b ? a : a + (double(x) / 2147483647.0) * (b - a);
  }
};

struct RunStats {
  int64_t loops = 0;
  int64_t convergences = 0;
  int64_t total_modifications = 0;
  int64_t evaluated_edges = 0;
  double seconds = 0.0;
  double shifted_energy = 0.0;  // sum of cluster energies after the run
};

template <class M>
struct Clustering {
  using Cluster = typename M::Cluster;
  const RefMesh& mesh;
  M metric;
  int K = 0;
  std::vector<int32_t> labels;
  std::vector<int32_t> sizes;
  std::vector<Cluster> clusters;
  std::vector<uint8_t> edge_last_loop;
  uint8_t rel_loops = 1;
  std::vector<int32_t> last_mod;
  std::deque<int64_t> queue;
  int connexity = 0;
  int loops = 0;
  int max_loops = 5000000;
  int max_convergences = 1000000000;
  int unconstrained_init = 0;
  int initial_sampling_type = 1;
  const int32_t* initial_labels = nullptr;
  std::vector<int32_t> fixed;  // FixedClusters: cluster k holds item fixed[k] (vtkUniformClustering.h:134)
  // IsClusterFreezed (vtkUniformClustering.h:329; set by the ForceManifold
  // restarts): edges touching a frozen cluster are skipped (:910-919) and
  // its items do not count for the early stop (:733-736); empty: none frozen
  std::vector<int32_t> frozen;
  bool is_frozen(int32_t c) const { return !frozen.empty() && frozen[size_t(c)] != 0; }
  int console = 0;
  int progress = 0;  // 1: one line per loop on stderr (long runs)
  RunStats stats;
  std::vector<EdgeTrace>* edge_trace = nullptr;
  size_t edge_trace_cap = 0;
  std::vector<LoopTrace>* loop_trace = nullptr;
  // energy log (SetWriteToGlobalEnergyLog): ComputeGlobalEnergy after each
  // loop (WriteToGlobalEnergyLog, :1383-1395) and after the run (:695-699)
  std::vector<long double>* energy_log = nullptr;
  long double final_energy = 0.0L;

  // ComputeGlobalEnergy (:1318-1346): the cluster energies summed in long
  // double over (int)sqrt(K) bins.  BinCount is never reset, so bin 0 takes
  // the first (int)sqrt(K) clusters and bin 1 every later one (:1333-1338).
  long double global_energy() const {
    const int nb = int(std::sqrt(double(K)));
    if (nb <= 0) return 0.0L;
    std::vector<long double> bins(size_t(nb), 0.0L);
    int bin = 0, count = 0;
    for (int i = 0; i < K; ++i) {
      bins[size_t(bin)] += (long double)M::energy(&clusters[i]);
      count++;
      if (count == nb && bin < nb - 1) bin++;
    }
    long double e = 0.0L;
    for (long double b : bins) e += b;
    return e;
  }
  int constrained_now = 1;  // the metric's ActiveConstraints (QEM), for the trace
  // scratch for the ring check
  std::vector<int32_t> vlist;
  std::deque<int32_t> vqueue;

  explicit Clustering(const RefMesh& m) : mesh(m) {}
  virtual ~Clustering() = default;

  int64_t n_items() const { return mesh.nv; }
  int64_t n_edges() const { return mesh.num_edges(); }

  void edge_items(int64_t e, int64_t& i1, int64_t& i2) const {
    i1 = mesh.ev1[e];
    i2 = mesh.ev2[e];
  }
  void push_ring(int64_t item) {
    for (int64_t k = mesh.ring_ptr[item]; k < mesh.ring_ptr[item + 1]; ++k) queue.push_back(mesh.ring[k]);
  }

  // vtkVerticesProcessing::ConnexityConstraintProblemLocal
  int connexity_problem(int64_t item, int64_t cluster) {
    if (!connexity) return 0;
    vlist.clear();
    vqueue.clear();
    for (int64_t k = mesh.ring_ptr[item]; k < mesh.ring_ptr[item + 1]; ++k) {
      int32_t e = mesh.ring[k];
      int32_t v1 = mesh.ev1[e], v2 = mesh.ev2[e];
      if (v1 != item && labels[v1] == cluster) vlist.push_back(v1);
      if (v2 != item && labels[v2] == cluster) vlist.push_back(v2);
    }
    if (vlist.empty()) return 0;
    size_t nverts = vlist.size();
    size_t visited = 1;
    auto delete_id = [&](int32_t id) {  // vtkIdList::DeleteId removes every occurrence
      std::vector<int32_t> keep;
      keep.reserve(vlist.size());
      for (int32_t x : vlist)
        if (x != id) keep.push_back(x);
      vlist.swap(keep);
    };
    int32_t v1 = vlist[0];
    delete_id(v1);
    vqueue.push_back(v1);
    while (!vqueue.empty()) {
      int32_t cur = vqueue.front();
      vqueue.pop_front();
      for (int64_t k = mesh.ring_ptr[cur]; k < mesh.ring_ptr[cur + 1]; ++k) {
        int32_t e = mesh.ring[k];
        int32_t v2 = mesh.ev1[e], v3 = mesh.ev2[e];
        for (size_t j = 0; j < vlist.size(); ++j) {
          int32_t x = vlist[j];
          if (x == v2) {
            vqueue.push_back(v2);
            visited++;
            delete_id(v2);
            break;
          } else if (x == v3) {
            vqueue.push_back(v3);
            visited++;
            delete_id(v3);
            break;
          }
        }
      }
    }
    return nverts == visited ? 0 : 1;
  }

  void set_number_of_clusters(int k) {
    K = k;
    clusters.assign(K, Cluster());
    for (auto& c : clusters) M::reset(&c);
  }

  void allocate() {
    sizes.assign(K, 0);
    labels.assign(n_items(), K);
    last_mod.assign(K, 0);
    edge_last_loop.assign(n_edges(), 0);
  }

  virtual void init(const double* indicator, double gradation) {
    allocate();
    metric.build(mesh, indicator, gradation);
    rel_loops = 1;
    loops = 0;
    queue.clear();
    for (int i = 0; i < K; ++i) sizes[i] = 0;
    for (auto& b : edge_last_loop) b = 0;
  }

  void recompute_sizes() {
    for (int i = 0; i < K; ++i) sizes[i] = 0;
    for (int64_t i = 0; i < n_items(); ++i) {
      int32_t c = labels[i];
      if (c >= 0 && c < K) sizes[c]++;
    }
  }

  void recompute_statistics() {
    recompute_sizes();
    for (int i = 0; i < K; ++i) M::reset(&clusters[i]);
    for (int64_t i = 0; i < n_items(); ++i) {
      int32_t c = labels[i];
      if (c >= 0 && c < K) metric.add_item(i, &clusters[c]);
    }
    for (int i = 0; i < K; ++i) {
      metric.compute_centroid(&clusters[i]);
      M::compute_energy(&clusters[i]);
    }
  }

  int clean_clustering() {
    std::vector<std::vector<int64_t>> comps(K);  // (first item, size) pairs, flattened
    // VisitedCluster[] (vtkUniformClustering.h:412, :463) is both the "seen"
    // flag and the first component's start item, so a component that starts
    // at item 0 is never recorded (and never reset).  Kept as is.
    std::vector<int64_t> first(K, 0), firstsize(K, 0);
    std::vector<uint8_t> visited(n_items(), 0);
    std::deque<int64_t> q;
    std::vector<int32_t> nb;
    int number = 0;
    for (int64_t i = 0; i < n_items(); ++i) {
      q.clear();
      if (visited[i] || labels[i] == K) continue;
      int64_t size = 0;
      q.push_back(i);
      int32_t type = labels[i];
      // is there a constraint on the item? (:436-447: the fixed item counts 1e9)
      const int64_t fixed_item = type < int32_t(fixed.size()) ? fixed[type] : -1;
      while (!q.empty()) {
        int64_t i1 = q.front();
        q.pop_front();
        if (visited[i1]) continue;
        size += i1 == fixed_item ? 1000000000 : 1;
        visited[i1] = 1;
        mesh.neighbours(i1, nb);
        for (int32_t i2 : nb)
          if (!visited[i2] && labels[i2] == type) q.push_back(i2);
      }
      if (first[type] == 0) {
        first[type] = i;
        firstsize[type] = size;
      } else {
        if (comps[type].empty()) {
          comps[type].push_back(first[type]);
          comps[type].push_back(firstsize[type]);
        }
        comps[type].push_back(i);
        comps[type].push_back(size);
      }
    }
    for (auto& v : visited) v = 0;
    for (int c = 0; c < K; ++c) {
      if (comps[c].empty()) continue;
      number++;
      int64_t sizemax = 0;
      size_t imax = 0;
      size_t nc = comps[c].size() / 2;
      for (size_t j = 0; j < nc; ++j) {
        if (sizemax >= comps[c][2 * j + 1]) continue;
        sizemax = comps[c][2 * j + 1];
        imax = j;
      }
      for (size_t j = 0; j < nc; ++j) {
        if (j == imax) continue;
        q.clear();
        q.push_back(comps[c][2 * j]);
        int32_t type = labels[comps[c][2 * j]];
        while (!q.empty()) {
          int64_t i1 = q.front();
          q.pop_front();
          if (visited[i1]) continue;
          visited[i1] = 1;
          labels[i1] = K;
          mesh.neighbours(i1, nb);
          for (int32_t i2 : nb)
            if (labels[i2] == type) q.push_back(i2);
        }
      }
    }
    return number;
  }

  void fill_holes() { fill_holes(labels); }

  // FillHolesInClustering on any label array whose null value is K (the
  // threaded layout reuses it on its region map).  The ring check always reads
  // the clustering itself, as in the reference.
  void fill_holes(std::vector<int32_t>& lab) {
    std::deque<int64_t> iq;
    for (int64_t e = 0; e < n_edges(); ++e) {
      int64_t i1, i2;
      edge_items(e, i1, i2);
      int32_t c1 = lab[i1], c2 = lab[i2];
      bool v1 = (c1 >= 0 && c1 < K), v2 = (c2 >= 0 && c2 < K);
      if (v1 != v2) iq.push_back(e);
    }
    while (!iq.empty()) {
      int64_t e = iq.front();
      iq.pop_front();
      int64_t i1, i2;
      edge_items(e, i1, i2);
      int32_t c1 = lab[i1], c2 = lab[i2];
      if (c1 == K) {
        std::swap(i1, i2);
        c1 = c2;
        c2 = K;
      }
      if (c1 != K && c2 == K && connexity_problem(i2, c2) == 0) {
        lab[i2] = c1;
        for (int64_t k = mesh.ring_ptr[i2]; k < mesh.ring_ptr[i2 + 1]; ++k) iq.push_back(mesh.ring[k]);
      }
    }
  }

  virtual void swap_queues() {}

  virtual void fill_queue() {
    queue.clear();
    for (int64_t e = 0; e < n_edges(); ++e) {
      int64_t i1, i2;
      edge_items(e, i1, i2);
      if (labels[i1] != labels[i2]) queue.push_back(e);
    }
    queue.push_back(-1);
  }

  void set_all_modified() {
    for (int i = 0; i < K; ++i) last_mod[i] = loops;
  }

  // vtkUniformClustering::ComputeInitialRandomSampling, metric weights, the
  // whole item range (List == 0); fixed item k starts in cluster k.
  void initial_random_sampling(std::vector<int32_t>& sampling, int nregions) {
    const int64_t n = n_items();
    sampling.assign(n, K);
    int offset = 0;
    for (; offset < int(fixed.size()); ++offset) sampling[fixed[offset]] = offset;
    std::vector<int32_t> items(n);
    int64_t remaining_items = n;
    int remaining_regions = nregions - offset;
    std::deque<int32_t> iq;
    std::vector<int32_t> nb;
    for (int64_t i = 0; i < n; ++i) items[i] = int32_t(i);
    std::mt19937 rng;
    rng.seed(0);
    for (int64_t i = n - 1; i > 0; --i) std::swap(items[i], items[rng() % uint64_t(n)]);
    double sw = 0;
    for (int64_t i = 0; i < n; ++i) sw += metric.item_weight(i);
    double weight = sw / double(nregions);
    int64_t first_item = 0;
    while (remaining_items > 0 && remaining_regions > 0) {
      bool found = false;
      int32_t item = 0;
      while (!found && first_item < n) {
        item = items[first_item];
        if (sampling[item] == K) found = true;
        else first_item++;
      }
      if (!found) break;  // every item labelled (the reference would spin here)
      iq.clear();
      iq.push_back(item);
      sw = 0;
      remaining_regions--;
      while (!iq.empty()) {
        item = iq.front();
        iq.pop_front();
        if (sampling[item] != K) continue;
        sampling[item] = remaining_regions + offset;
        sw += metric.item_weight(item);
        remaining_items--;
        mesh.neighbours(item, nb);
        for (int32_t j : nb) iq.push_back(j);
        if (sw > weight) break;
      }
    }
    if (remaining_regions == 0) return;
    // more regions than reachable items: reassign single items
    // (vtkUniformClustering.h:1269-1311)
    std::vector<int32_t> csz(size_t(std::max(K, nregions)) + 1, 0);
    for (int64_t i = 0; i < n; ++i) {
      items[i] = int32_t(i);
      int32_t c = sampling[i];
      if (c == K) continue;
      csz[c]++;
    }
    for (int64_t i = n - 1; i > 0; --i) std::swap(items[i], items[rng() % uint64_t(n)]);
    first_item = 0;
    while (remaining_regions) {
      int32_t item = 0;
      while (true) {
        item = items[first_item];
        first_item++;
        int32_t c = sampling[item];
        if (c == K) break;
        if (csz[c] == 1) continue;
        sampling[item] = remaining_regions + offset;
        csz[c]--;
        csz[remaining_regions + offset]++;
        break;
      }
      remaining_regions--;
      sampling[item] = remaining_regions + offset;
    }
  }

  void init_samples(uint64_t type0_seed = 5000) {
    for (int64_t i = 0; i < n_items(); ++i) labels[i] = K;
    if (initial_sampling_type == 2 && initial_labels) {
      for (int64_t i = 0; i < n_items(); ++i) labels[i] = initial_labels[i];
    } else if (initial_sampling_type == 1) {
      initial_random_sampling(labels, K);
    } else {
      // type 0 (:1126-1159): vtkMath::RandomSeed(5000), one Random(), then
      // per cluster Random(0, n) as float, rounded, redrawn while taken (a
      // draw rounding to n, past the last item, is redrawn as well)
      VtkMinStd rng;
      rng.set_seed(int32_t(type0_seed));
      (void)rng.uniform();
      const float size = float(n_items());
      for (int c = 0; c < K; ++c) {
        for (;;) {
          const float test = float(rng.range(0.0, double(size)));
          const int64_t number = int64_t(std::floor(double(test) + 0.5));
          if (number >= n_items()) continue;
          if (labels[number] == K) {
            labels[number] = c;
            break;
          }
        }
      }
    }
  }

  virtual int process_one_loop() {
    Cluster c21, c22, c31, c32;
    int mods = 0;
    while (true) {
      int64_t edge = queue.front();
      queue.pop_front();
      if (edge == -1) {
        queue.push_back(-1);
        return mods;
      }
      int64_t i1, i2;
      edge_items(edge, i1, i2);
      if (edge_last_loop[edge] == rel_loops || i2 < 0) continue;
      edge_last_loop[edge] = rel_loops;
      int32_t v1 = labels[i1], v2 = labels[i2];
      if (v1 == v2) continue;
      if (v1 == K) {
        Cluster* c2 = &clusters[v2];
        metric.add_item(i1, c2);
        metric.compute_centroid(c2);
        M::compute_energy(c2);
        sizes[v2]++;
        push_ring(i1);
        mods++;
        labels[i1] = v2;
        last_mod[v2] = loops;
        continue;
      } else if (v2 == K) {
        Cluster* c1 = &clusters[v1];
        metric.add_item(i2, c1);
        metric.compute_centroid(c1);
        M::compute_energy(c1);
        sizes[v1]++;
        push_ring(i2);
        mods++;
        labels[i2] = v1;
        last_mod[v1] = loops;
        continue;
      }
      if ((last_mod[v1] < loops - 1 && last_mod[v2] < loops - 1) || is_frozen(v1) || is_frozen(v2)) {
        queue.push_back(edge);
        continue;
      }
      stats.evaluated_edges++;
      Cluster* c1 = &clusters[v1];
      Cluster* c2 = &clusters[v2];
      double try1 = M::energy(c1) + M::energy(c2);
      double try2, try3;
      if (sizes[v1] == 1 || connexity_problem(i1, v1) == 1) {
        try2 = 100000000.0;
      } else {
        M::copy(c1, &c21);
        metric.sub_item(i1, &c21);
        M::copy(c2, &c22);
        metric.add_item(i1, &c22);
        metric.compute_centroid(&c21);
        metric.compute_centroid(&c22);
        M::compute_energy(&c21);
        M::compute_energy(&c22);
        try2 = M::energy(&c21) + M::energy(&c22);
      }
      if (sizes[v2] == 1 || connexity_problem(i2, v2) == 1) {
        try3 = 1000000000.0;
      } else {
        M::copy(c2, &c32);
        metric.sub_item(i2, &c32);
        M::copy(c1, &c31);
        metric.add_item(i2, &c31);
        metric.compute_centroid(&c31);
        metric.compute_centroid(&c32);
        M::compute_energy(&c31);
        M::compute_energy(&c32);
        try3 = M::energy(&c31) + M::energy(&c32);
      }
      // stay when Try1 is not above either move; else the strictly lower
      // of Try2 (move I1) and Try3 (move I2), a tie moving I2
      const int32_t decision = (try1 <= try2 && try1 <= try3) ? 0 : (try2 < try1 && try2 < try3) ? 1 : 2;
      if (edge_trace && edge_trace->size() < edge_trace_cap)
        edge_trace->push_back(EdgeTrace{loops, decision, edge, i1, i2, try1, try2, try3});
      if (decision == 0) {
        queue.push_back(edge);
      } else if (decision == 1) {
        labels[i1] = v2;
        sizes[v2]++;
        sizes[v1]--;
        M::copy(&c21, c1);
        M::copy(&c22, c2);
        push_ring(i1);
        mods++;
        last_mod[v1] = loops;
        last_mod[v2] = loops;
      } else {
        labels[i2] = v1;
        sizes[v1]++;
        sizes[v2]--;
        M::copy(&c31, c1);
        M::copy(&c32, c2);
        push_ring(i2);
        mods++;
        last_mod[v1] = loops;
        last_mod[v2] = loops;
      }
    }
  }

  void minimize_energy() {
    fill_holes();
    fill_queue();
    recompute_statistics();
    set_all_modified();
    int convergences = 0;
    int64_t early_stop_items = 0;
    for (int i = 0; i < K; ++i)
      if (!is_frozen(i)) early_stop_items += sizes[i];
    while (true) {
      swap_queues();
      const int loop_connexity = connexity;
      int mods = process_one_loop();
      if (energy_log) energy_log->push_back(global_energy());
      stats.total_modifications += mods;
      if (progress) std::fprintf(stderr, "%sLoop %d: %d Modifications\n", connexity ? "*" : "", loops, mods);
      LoopTrace lt{loops, mods, loop_connexity, constrained_now, 0, 0};
      loops++;
      if (rel_loops == 255) {
        for (auto& b : edge_last_loop) b = 0;
        rel_loops = 1;
      } else {
        rel_loops++;
      }
      const bool conv = mods == 0 || loops > max_loops || (mods <= early_stop_items / 1000 && convergences <= 1);
      if (!conv && loop_trace) loop_trace->push_back(lt);
      if (conv) {
        if (unconstrained_init && convergences == 0) {
          metric.set_constrained(1);
          constrained_now = 1;
        }
        if (convergences >= 1) connexity = 1;
        convergences++;
        int ndisc = clean_clustering();
        fill_holes();
        recompute_sizes();
        if (loop_trace) {
          lt.convergence = 1;
          lt.disconnected = ndisc;
          loop_trace->push_back(lt);
        }
        if (ndisc == 0 && mods == 0) break;
        if (loops >= max_loops) break;
        if (convergences >= max_convergences) break;
        recompute_statistics();
        fill_queue();
        set_all_modified();
      }
    }
    stats.loops = loops;
    stats.convergences = convergences;
  }

  // ProcessClustering: Init, InitSamples, optional unconstrained start,
  // MinimizeEnergy (timed, like the reference's timer around MinimizeEnergy).
  void process(const double* indicator, double gradation) {
    init(indicator, gradation);
    init_samples();
    if (unconstrained_init) {
      metric.set_constrained(0);
      constrained_now = 0;
    }
    auto t0 = std::chrono::steady_clock::now();
    minimize_energy();
    auto t1 = std::chrono::steady_clock::now();
    stats.seconds = std::chrono::duration<double>(t1 - t0).count();
    if (energy_log) {
      recompute_statistics();
      final_energy = global_energy();
    }
    double e = 0;
    for (int i = 0; i < K; ++i) e += M::energy(&clusters[i]);
    stats.shifted_energy = e;
  }
};

}  // namespace oracle
