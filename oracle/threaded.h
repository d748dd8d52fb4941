// TEST INFRASTRUCTURE — CPU oracle, not part of the product.
//
// Oracle B: restatement of the reference's multithreaded engine
// DiscreteRemeshing/vtkThreadedClustering.h (ACVDP / ACVDQP, compiled in by
// `#define DOmultithread`, DiscreteRemeshing/Examples/ACVDP.cxx:1-3):
//   Init                   :773-804   PoolSize = PoolingRatio * T + 1
//   ComputeEdgesLayout     :806-888   type 1: PoolSize-1 regions from the
//                                      reference initial sampling + fill holes;
//                                      an edge belongs to its region when both
//                                      ends are in it, else to the last
//                                      (serial, cross-region) process
//   FillQueuesFromClustering :204-242
//   MyMainForClustering    :245-283   threads pop processes from a priority
//                                      queue (longest last duration first)
//   ExecuteProcess         :285-493   the serial decision rule with per-cluster
//                                      locks taken in id order; null absorption
//                                      does not stamp the clusters
//   ProcessOneLoop         :662-693   T threads, then the cross-region process
//                                      on the calling thread
//   SwapQueues             :765-770
// It is the timed CPU baseline.  Like the reference it is non-deterministic
// (labels are read without locks by the ring check and across regions).
#pragma once

#include <atomic>
#include <chrono>
#include <mutex>
#include <queue>
#include <thread>
#include <vector>

#include "serial.h"

namespace oracle {

// vtkPriorityQueue subset: Insert ignores ids already queued, Pop returns the
// id of smallest priority (-1 when empty).
class IdPriorityQueue {
 public:
  void reset(int n) {
    present_.assign(n, 0);
    heap_ = decltype(heap_)();
  }
  void insert(double prio, int id) {
    if (present_[id]) return;
    present_[id] = 1;
    heap_.emplace(prio, id);
  }
  int pop() {
    if (heap_.empty()) return -1;
    int id = heap_.top().second;
    heap_.pop();
    present_[id] = 0;
    return id;
  }

 private:
  std::vector<char> present_;
  std::priority_queue<std::pair<double, int>, std::vector<std::pair<double, int>>, std::greater<>> heap_;
};

template <class M>
struct ThreadedClustering : public Clustering<M> {
  using Base = Clustering<M>;
  using Cluster = typename M::Cluster;
  int num_threads = 1;
  int pooling_ratio = 5;
  int pool_size = 0;
  std::vector<int32_t> edges_process;
  // queues[dest][src]
  std::vector<std::vector<std::queue<int32_t>>> q1, q2;
  std::vector<std::vector<std::queue<int32_t>>>* push = nullptr;
  std::vector<std::vector<std::queue<int32_t>>>* pop = nullptr;
  std::vector<std::mutex> locks;
  std::mutex pool_lock, pool_lock2;
  IdPriorityQueue pq_a, pq_b;
  IdPriorityQueue* pq1 = &pq_a;
  IdPriorityQueue* pq2 = &pq_b;
  std::vector<int64_t> thread_mods;
  int64_t locking_collisions = 0;

  explicit ThreadedClustering(const RefMesh& m) : Base(m) {}

  void init(const double* indicator, double gradation) override {
    Base::init(indicator, gradation);
    pool_size = pooling_ratio * num_threads + 1;
    compute_edges_layout();
    q1.assign(pool_size, std::vector<std::queue<int32_t>>(pool_size));
    q2.assign(pool_size, std::vector<std::queue<int32_t>>(pool_size));
    push = &q1;
    pop = &q2;
    pq_a.reset(pool_size);
    pq_b.reset(pool_size);
    locks = std::vector<std::mutex>(this->K + 1);
    thread_mods.assign(num_threads + 1, 0);
  }

  void compute_edges_layout() {
    std::vector<int32_t> temp;
    this->initial_random_sampling(temp, pool_size - 1);
    this->fill_holes(temp);
    const int64_t ne = this->n_edges();
    edges_process.assign(ne, pool_size - 1);
    for (int64_t e = 0; e < ne; ++e) {
      int64_t i1, i2;
      this->edge_items(e, i1, i2);
      int32_t c1 = temp[i1];
      if (c1 >= pool_size || c1 < 0) c1 = pool_size - 1;
      if (temp[i1] != c1) c1 = pool_size - 1;
      if (temp[i2] != c1) c1 = pool_size - 1;
      edges_process[e] = c1;
    }
  }

  void fill_queue() override {
    for (auto& row : *push)
      for (auto& q : row)
        while (!q.empty()) q.pop();
    for (int64_t e = 0; e < this->n_edges(); ++e) {
      int64_t i1, i2;
      this->edge_items(e, i1, i2);
      if (this->labels[i1] != this->labels[i2]) {
        int p = edges_process[e];
        (*push)[p][p].push(int32_t(e));
      }
    }
    for (int i = 0; i < pool_size - 1; ++i) pq1->insert(double(i), i);
  }

  void swap_queues() override { std::swap(push, pop); }

  void add_edge(int32_t e, int process) { (*push)[edges_process[e]][process].push(e); }

  void add_ring(int64_t item, int process) {
    const RefMesh& m = this->mesh;
    for (int64_t k = m.ring_ptr[item]; k < m.ring_ptr[item + 1]; ++k) add_edge(m.ring[k], process);
  }

  // Thread-local ring-check scratch (the reference keeps one VList/VQueue per
  // thread, vtkVerticesProcessing.h:100-128).
  struct Scratch {
    std::vector<int32_t> vlist;
    std::deque<int32_t> vqueue;
  };

  int connexity_problem_local(int64_t item, int64_t cluster, Scratch& s) {
    if (!this->connexity) return 0;
    const RefMesh& m = this->mesh;
    const std::vector<int32_t>& labels = this->labels;
    s.vlist.clear();
    s.vqueue.clear();
    for (int64_t k = m.ring_ptr[item]; k < m.ring_ptr[item + 1]; ++k) {
      int32_t e = m.ring[k];
      int32_t v1 = m.ev1[e], v2 = m.ev2[e];
      if (v1 != item && labels[v1] == cluster) s.vlist.push_back(v1);
      if (v2 != item && labels[v2] == cluster) s.vlist.push_back(v2);
    }
    if (s.vlist.empty()) return 0;
    size_t nverts = s.vlist.size(), visited = 1;
    auto delete_id = [&](int32_t id) {
      size_t w = 0;
      for (size_t r = 0; r < s.vlist.size(); ++r)
        if (s.vlist[r] != id) s.vlist[w++] = s.vlist[r];
      s.vlist.resize(w);
    };
    int32_t v1 = s.vlist[0];
    delete_id(v1);
    s.vqueue.push_back(v1);
    while (!s.vqueue.empty()) {
      int32_t cur = s.vqueue.front();
      s.vqueue.pop_front();
      for (int64_t k = m.ring_ptr[cur]; k < m.ring_ptr[cur + 1]; ++k) {
        int32_t e = m.ring[k];
        int32_t v2 = m.ev1[e], v3 = m.ev2[e];
        for (size_t j = 0; j < s.vlist.size(); ++j) {
          int32_t x = s.vlist[j];
          if (x == v2) {
            s.vqueue.push_back(v2);
            visited++;
            delete_id(v2);
            break;
          } else if (x == v3) {
            s.vqueue.push_back(v3);
            visited++;
            delete_id(v3);
            break;
          }
        }
      }
    }
    return nverts == visited ? 0 : 1;
  }

  void lock_pair(int32_t a, int32_t b) {
    int32_t lo = a < b ? a : b, hi = a < b ? b : a;
    if (!locks[lo].try_lock()) locks[lo].lock();
    if (!locks[hi].try_lock()) locks[hi].lock();
  }

  void execute_process(int process, int thread) {
    Cluster c21, c22, c31, c32;
    Scratch scratch;
    const int K = this->K;
    auto& labels = this->labels;
    auto& sizes = this->sizes;
    auto& metric = this->metric;
    for (int cq = 0; cq < pool_size; ++cq) {
      std::queue<int32_t>& queue = (*pop)[process][cq];
      while (!queue.empty()) {
        int32_t edge = queue.front();
        queue.pop();
        int64_t i1, i2;
        this->edge_items(edge, i1, i2);
        if (this->edge_last_loop[edge] == this->rel_loops || i2 < 0) continue;
        this->edge_last_loop[edge] = this->rel_loops;
        int32_t v1 = labels[i1], v2 = labels[i2];
        if (v1 == v2) continue;
        lock_pair(v1, v2);
        if (v1 == K) {
          Cluster* c = &this->clusters[v2];
          metric.add_item(i1, c);
          metric.compute_centroid(c);
          M::compute_energy(c);
          sizes[v2]++;
          add_ring(i1, process);
          labels[i1] = v2;
          thread_mods[thread]++;
        } else if (v2 == K) {
          Cluster* c = &this->clusters[v1];
          metric.add_item(i2, c);
          metric.compute_centroid(c);
          M::compute_energy(c);
          sizes[v1]++;
          add_ring(i2, process);
          labels[i2] = v1;
          thread_mods[thread]++;
        } else {
          int result = 1;
          Cluster* c1 = &this->clusters[v1];
          Cluster* c2 = &this->clusters[v2];
          if (this->last_mod[v1] >= this->loops - 1 || this->last_mod[v2] >= this->loops - 1) {
            double try1 = M::energy(c1) + M::energy(c2), try2, try3;
            if (sizes[v1] == 1 || connexity_problem_local(i1, v1, scratch) == 1) {
              try2 = 100000000.0;
            } else {
              M::copy(c1, &c21);
              M::copy(c2, &c22);
              metric.sub_item(i1, &c21);
              metric.add_item(i1, &c22);
              metric.compute_centroid(&c21);
              metric.compute_centroid(&c22);
              M::compute_energy(&c21);
              M::compute_energy(&c22);
              try2 = M::energy(&c21) + M::energy(&c22);
            }
            if (sizes[v2] == 1 || connexity_problem_local(i2, v2, scratch) == 1) {
              try3 = 1000000000.0;
            } else {
              M::copy(c1, &c31);
              M::copy(c2, &c32);
              metric.sub_item(i2, &c32);
              metric.add_item(i2, &c31);
              metric.compute_centroid(&c31);
              metric.compute_centroid(&c32);
              M::compute_energy(&c31);
              M::compute_energy(&c32);
              try3 = M::energy(&c31) + M::energy(&c32);
            }
            if (try1 <= try2 && try1 <= try3) result = 1;
            else if (try2 < try1 && try2 < try3) result = 2;
            else result = 3;
          }
          if (result == 1) {
            add_edge(edge, process);
          } else if (result == 2) {
            labels[i1] = v2;
            sizes[v2]++;
            sizes[v1]--;
            M::copy(&c21, c1);
            M::copy(&c22, c2);
            add_ring(i1, process);
            thread_mods[thread]++;
            this->last_mod[v1] = this->loops;
            this->last_mod[v2] = this->loops;
          } else {
            labels[i2] = v1;
            sizes[v1]++;
            sizes[v2]--;
            M::copy(&c31, c1);
            M::copy(&c32, c2);
            add_ring(i2, process);
            thread_mods[thread]++;
            this->last_mod[v1] = this->loops;
            this->last_mod[v2] = this->loops;
          }
        }
        locks[v1].unlock();
        locks[v2].unlock();
      }
    }
  }

  int process_one_loop() override {
    for (auto& t : thread_mods) t = 0;
    auto worker = [this](int tid) {
      for (;;) {
        int process;
        {
          std::lock_guard<std::mutex> g(pool_lock);
          process = pq1->pop();
        }
        if (process < 0) break;
        auto t0 = std::chrono::steady_clock::now();
        execute_process(process, tid);
        double dt = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        std::lock_guard<std::mutex> g(pool_lock2);
        pq2->insert(-dt, process);
      }
    };
    // vtkMultiThreader::SingleMethodExecute: T-1 spawned threads + the caller
    std::vector<std::thread> th;
    for (int t = 1; t < num_threads; ++t) th.emplace_back(worker, t);
    worker(0);
    for (auto& t : th) t.join();
    execute_process(pool_size - 1, num_threads);
    std::swap(pq1, pq2);
    int64_t mods = 0;
    for (auto v : thread_mods) mods += v;
    return int(mods);
  }
};

}  // namespace oracle
