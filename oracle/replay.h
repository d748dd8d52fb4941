// TEST INFRASTRUCTURE — CPU oracle, not part of the product.
//
// Replay C: the GPU's pass schedule executed sequentially on the host for one
// mesh.  It reads the same inputs the device sweeps (CSR adjacency, FP32
// items, fixed-point scales) and calls the same per-item rule and cluster
// arithmetic (acvd_amd/csrc/acvd_math.h, compiled here with
// -ffp-contract=off), so its final labels must equal the device's bit for bit.
// What it restates independently of the device code: the pass structure
// (propose over a frozen state, rounds of min-key accept and exclusive apply
// with re-evaluation of the proposals left), the
// synchronous fill rounds, the largest-component clean and the host-side
// phase decisions between convergences.
#pragma once

#include <cstdint>
#include <deque>
#include <vector>

#include "../acvd_amd/csrc/acvd_math.h"

namespace oracle {

struct ReplayStats {
  int64_t loops = 0, convergences = 0, total_moves = 0, evaluated = 0, proposed = 0;
  int32_t last_disconnected = 0;
  // moves to a cluster no neighbour of the item belongs to (the reference
  // moves items across cut edges only: must stay 0)
  int64_t non_adjacent_moves = 0;
};

struct Replay {
  int64_t N = 0;
  int32_t K = 0;
  int metric = 0, level = 3;
  uint32_t seed = 0;
  acvd::Scales sc{};
  int max_loops = 5000000, max_conv = 1000000000, unconstrained_init = 0;
  bool generic_qem = false;  // force the generic QEM proposal path (equivalence checks)
  bool all_rounds = false;   // run every round even after one without moves (equivalence checks)
  int first_rounds = 0;  // accept rounds per pass before the first convergence (0: the library's rule)
  const int32_t* rp = nullptr;
  const int32_t* col = nullptr;
  const float* xw = nullptr;
  const float* q = nullptr;

  std::vector<int32_t> labels, sizes, stamp, cand_pass, pending, cl_round;
  std::vector<int32_t> anchor;  // per cluster: its fixed item or -1 (empty: no fixed clusters)
  std::vector<int32_t> frozen;  // per cluster: nonzero when frozen (empty: none)
  std::vector<unsigned long long> acc_ws, acc_q, ring_adj;
  std::vector<acvd::ClusterGeom> geom;
  std::vector<double> T, evec;
  std::vector<unsigned long long> prio;
  acvd::MeshState st{};
  ReplayStats stats;
  // energy log (the library's write_energy_log): the energy after each pass
  // and after the run, summed as the device sums it
  bool log_energy = false;
  std::vector<double> energy_log;
  double final_energy = 0.0;

  // Energy of the current labels: each cluster's log_cluster_energy from its
  // accumulators with a cold solve, summed in the fixed order of kLogThreads
  // (acvd_math.h).
  double log_sum() const {
    double part[acvd::kLogThreads];
    for (int t = 0; t < acvd::kLogThreads; ++t) part[t] = 0.0;
    for (int32_t c = 0; c < K; ++c) {
      const int64_t* ws = reinterpret_cast<const int64_t*>(&acc_ws[4 * int64_t(c)]);
      const acvd::ClusterGeom g = acvd::geom_from_fixed(ws[0], ws[1], ws[2], ws[3], sc);
      double Tc = 0.0;
      if (metric == 1) {
        if (!anchor.empty() && anchor[c] >= 0) {
          const acvd::AnchorPos a{double(xw[4 * int64_t(anchor[c])]), double(xw[4 * int64_t(anchor[c]) + 1]),
                                  double(xw[4 * int64_t(anchor[c]) + 2]), true};
          Tc = acvd::anchored_T(g, a);
        } else if (st.constrained) {
          Tc = acvd::qem_T_from_fixed(reinterpret_cast<const int64_t*>(&acc_q[10 * int64_t(c)]), g, sc, level,
                                      nullptr);
        }
      }
      part[c % acvd::kLogThreads] += acvd::log_cluster_energy(g, Tc);
    }
    for (int h = acvd::kLogThreads / 2; h > 0; h >>= 1)
      for (int t = 0; t < h; ++t) part[t] += part[t + h];
    return part[0];
  }

  acvd::SweepView view() const {
    acvd::SweepView v;
    v.row_ptr = rp;
    v.col = col;
    v.labels = labels.data();
    v.xw = xw;
    v.q = q;
    v.acc_ws = acc_ws.data();
    v.acc_q = acc_q.data();
    v.evec = metric == 1 ? evec.data() : nullptr;
    v.geom = geom.data();
    v.T = T.data();
    v.sizes = sizes.data();
    v.stamp = stamp.data();
    v.cand_pass = cand_pass.data();
    v.ring_adj = ring_adj.empty() ? nullptr : ring_adj.data();
    v.anchor = anchor.empty() ? nullptr : anchor.data();
    v.frozen = frozen.empty() ? nullptr : frozen.data();
    return v;
  }

  void fill_until_stable() {
    for (;;) {
      bool changed = false;
      for (int64_t i = 0; i < N; ++i) {
        if (labels[i] != acvd::kNull) continue;
        int32_t pick = acvd::kNull;
        for (int32_t k = rp[i]; k < rp[i + 1]; ++k)
          if (labels[col[k]] != acvd::kNull) {
            pick = labels[col[k]];
            break;
          }
        pending[i] = pick;
        changed |= pick != acvd::kNull;
      }
      for (int64_t i = 0; i < N; ++i)
        if (labels[i] == acvd::kNull) labels[i] = pending[i];
      if (!changed) break;
    }
  }

  // warm: the derive after a move, warm-started from the cluster's last
  // eigenvectors (as on the device); the recompute solves cold
  void derive(int32_t c, bool warm) {
    const int64_t* ws = reinterpret_cast<const int64_t*>(&acc_ws[4 * int64_t(c)]);
    geom[c] = acvd::geom_from_fixed(ws[0], ws[1], ws[2], ws[3], sc);
    if (metric == 1) {
      const int64_t* q10 = reinterpret_cast<const int64_t*>(&acc_q[10 * int64_t(c)]);
      double V0[9];
      for (int k = 0; k < 9; ++k) V0[k] = evec[9 * int64_t(c) + k];
      if (!anchor.empty() && anchor[c] >= 0) {
        const acvd::AnchorPos a{double(xw[4 * int64_t(anchor[c])]), double(xw[4 * int64_t(anchor[c]) + 1]),
                                double(xw[4 * int64_t(anchor[c]) + 2]), true};
        T[c] = acvd::anchored_T(geom[c], a);
      } else {
        T[c] = st.constrained ? acvd::qem_T_from_fixed(q10, geom[c], sc, level, nullptr, warm ? V0 : nullptr,
                                                       &evec[9 * int64_t(c)])
                              : 0.0;
      }
    }
  }

  int64_t recompute(bool set_stamp) {
    std::fill(acc_ws.begin(), acc_ws.end(), 0ull);
    std::fill(acc_q.begin(), acc_q.end(), 0ull);
    std::fill(sizes.begin(), sizes.end(), 0);
    int64_t count = 0;
    for (int64_t i = 0; i < N; ++i) {
      const int32_t c = labels[i];
      if (c == acvd::kNull) continue;
      const float w = xw[4 * i + 3];
      acc_ws[4 * int64_t(c)] += (unsigned long long)acvd::fix_w(w, sc);
      for (int k = 0; k < 3; ++k)
        acc_ws[4 * int64_t(c) + 1 + k] += (unsigned long long)acvd::fix_s(w, xw[4 * i + k], sc);
      if (metric == 1)
        for (int k = 0; k < 10; ++k) acc_q[10 * int64_t(c) + k] += (unsigned long long)acvd::fix_q(q[10 * i + k], sc);
      sizes[c]++;
      if (frozen.empty() || !frozen[size_t(c)]) count++;  // early-stop items
    }
    for (int32_t c = 0; c < K; ++c) {
      derive(c, false);
      if (set_stamp) stamp[c] = st.loops;
    }
    return count;
  }

  void update(int32_t c, int sign, int64_t iw, const int64_t* is, const int64_t* iq) {
    acc_ws[4 * int64_t(c)] += (unsigned long long)(sign * iw);
    for (int k = 0; k < 3; ++k) acc_ws[4 * int64_t(c) + 1 + k] += (unsigned long long)(sign * is[k]);
    if (metric == 1)
      for (int k = 0; k < 10; ++k) acc_q[10 * int64_t(c) + k] += (unsigned long long)(sign * iq[k]);
    derive(c, true);
  }

  void apply_move(int64_t i, int32_t lab, int32_t b, int32_t L, int32_t rtag) {
    bool adjacent = false;
    for (int32_t k = rp[i]; k < rp[i + 1] && !adjacent; ++k) adjacent = labels[col[k]] == b;
    if (!adjacent) stats.non_adjacent_moves++;
    labels[i] = b;
    const float w = xw[4 * i + 3];
    const int64_t iw = acvd::fix_w(w, sc);
    const int64_t is[3] = {acvd::fix_s(w, xw[4 * i], sc), acvd::fix_s(w, xw[4 * i + 1], sc),
                           acvd::fix_s(w, xw[4 * i + 2], sc)};
    int64_t iq[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    if (metric == 1)
      for (int k = 0; k < 10; ++k) iq[k] = acvd::fix_q(q[10 * i + k], sc);
    if (lab != acvd::kNull) {
      update(lab, -1, iw, is, iq);
      sizes[lab]--;
      stamp[lab] = L;
      cl_round[lab] = rtag;
    }
    update(b, +1, iw, is, iq);
    sizes[b]++;
    stamp[b] = L;
    cl_round[b] = rtag;
  }

  bool propose(int64_t i, int32_t* target) {
    const acvd::SweepView v = view();
    const int32_t lab = labels[i];
    if (metric == 0) return acvd::propose_item<0>(v, i, lab, st.connexity, st.constrained, level, sc, target);
    const int mode = acvd::qem_mode(st);
    return generic_qem ? acvd::propose_item<1, false>(v, i, lab, st.connexity, mode, level, sc, target)
                       : acvd::propose_item<1>(v, i, lab, st.connexity, mode, level, sc, target);
  }

  void pass() {
    const int32_t L = st.loops;
    const acvd::SweepView v = view();
    // round 0 proposals: every item needs_eval selects, against the frozen state
    std::vector<int64_t> cand;
    std::vector<int32_t> target;
    for (int64_t i = 0; i < N; ++i) {
      if (!acvd::needs_eval(v, i, labels[i], L)) continue;
      stats.evaluated++;
      int32_t t = acvd::kNull;
      if (!propose(i, &t)) continue;
      cand.push_back(i);
      target.push_back(t);
    }
    stats.proposed += int64_t(cand.size());
    for (int64_t i : cand) cand_pass[i] = L;
    int mods = 0;
    std::vector<int32_t> from(cand.size());
    std::vector<int32_t> since(cand.size(), L * acvd::kMatchRounds);
    const int nrounds = acvd::pass_rounds(st);
    for (int round = 0; round < nrounds; ++round) {
      const int mods_before = mods;
      if (round > 0) {
        // a live proposal whose own cluster or target moved since its last
        // evaluation is evaluated again against the state left by the last
        // round; the others keep their (still exact) move
        for (size_t t = 0; t < cand.size(); ++t) {
          if (target[t] < 0) continue;
          const int32_t lab = labels[cand[t]];
          const bool stale = (lab != acvd::kNull && cl_round[lab] >= since[t]) || cl_round[target[t]] >= since[t];
          if (!stale) continue;
          since[t] = L * acvd::kMatchRounds + round;
          if (lab != acvd::kNull && (metric == 0 || acvd::qem_mode(st) == 0)) {
            // the isotropic rule: the move itself is checked again
            target[t] = acvd::iso_recheck(view(), cand[t], lab, target[t], st.connexity) ? target[t] : -3;
            continue;
          }
          int32_t nt = acvd::kNull;
          target[t] = propose(cand[t], &nt) ? nt : -3;
        }
        // (all re-evaluations of a round see the same state: none is applied before
        // the keys below are compared)
      }
      for (size_t t = 0; t < cand.size(); ++t) from[t] = labels[cand[t]];
      for (size_t t = 0; t < cand.size(); ++t) {
        if (target[t] < 0) continue;
        const uint64_t key = acvd::accept_key(seed, L, round, uint32_t(cand[t]));
        if (from[t] != acvd::kNull && key < prio[from[t]]) prio[from[t]] = key;
        if (key < prio[target[t]]) prio[target[t]] = key;
      }
      for (size_t t = 0; t < cand.size(); ++t) {
        if (target[t] < 0) continue;
        const uint64_t key = acvd::accept_key(seed, L, round, uint32_t(cand[t]));
        if (from[t] != acvd::kNull && prio[from[t]] != key) continue;
        if (prio[target[t]] != key) continue;
        apply_move(cand[t], from[t], target[t], L, L * acvd::kMatchRounds + round);
        target[t] = -3;
        mods++;
      }
      // a round without moves ends the pass (later rounds would apply nothing)
      if (!all_rounds && mods == mods_before) break;
    }
    st.pass_mods = mods;
    st.pass_props = int32_t(cand.size());
  }

  // Largest connected component per cluster (ties: smallest first item);
  // other components become null.  Returns the number of clusters that had
  // more than one component.
  int clean() {
    std::vector<int32_t> comp(N, -1);
    std::vector<int64_t> comp_size;
    std::vector<int64_t> comp_root;
    std::deque<int64_t> q2;
    for (int64_t i = 0; i < N; ++i) {
      if (comp[i] >= 0 || labels[i] == acvd::kNull) continue;
      const int32_t id = int32_t(comp_size.size());
      const int32_t lab = labels[i];
      int64_t size = 0;
      comp[i] = id;
      q2.clear();
      q2.push_back(i);
      while (!q2.empty()) {
        int64_t u = q2.front();
        q2.pop_front();
        size++;
        for (int32_t k = rp[u]; k < rp[u + 1]; ++k) {
          int32_t v = col[k];
          if (comp[v] < 0 && labels[v] == lab) {
            comp[v] = id;
            q2.push_back(v);
          }
        }
      }
      comp_size.push_back(size);
      comp_root.push_back(i);
    }
    // the component holding a cluster's fixed item is kept (its size counts
    // 1e9 in CleanClustering, vtkUniformClustering.h:441-447)
    if (!anchor.empty())
      for (int32_t c = 0; c < K; ++c)
        if (anchor[c] >= 0 && labels[anchor[c]] == c) comp_size[comp[anchor[c]]] += int64_t(1) << 40;
    std::vector<int32_t> best(K, -1), ncomp(K, 0);
    for (size_t c = 0; c < comp_size.size(); ++c) {
      const int32_t lab = labels[comp_root[c]];
      ncomp[lab]++;
      const int32_t b = best[lab];
      if (b < 0 || comp_size[c] > comp_size[b]) best[lab] = int32_t(c);  // roots increase: first max kept
    }
    for (int64_t i = 0; i < N; ++i) {
      const int32_t lab = labels[i];
      if (lab == acvd::kNull) continue;
      if (comp[i] != best[lab]) labels[i] = acvd::kNull;
    }
    int nd = 0;
    for (int32_t c = 0; c < K; ++c) nd += ncomp[c] > 1;
    return nd;
  }

  // init: labels in the reference convention (0..K-1, K = null)
  void run(const int32_t* init) {
    ring_adj.resize(N);
    for (int64_t i = 0; i < N; ++i) ring_adj[i] = acvd::ring_adjacency(rp, col, int32_t(i));
    labels.assign(N, acvd::kNull);
    for (int64_t i = 0; i < N; ++i) labels[i] = (init[i] < 0 || init[i] >= K) ? acvd::kNull : init[i];
    sizes.assign(K, 0);
    stamp.assign(K, 0);
    cand_pass.assign(N, -2);
    cl_round.assign(K, -1);
    pending.assign(N, acvd::kNull);
    acc_ws.assign(4 * size_t(K), 0ull);
    acc_q.assign(metric == 1 ? 10 * size_t(K) : 0, 0ull);
    geom.assign(K, acvd::ClusterGeom{});
    T.assign(metric == 1 ? K : 0, 0.0);
    evec.assign(metric == 1 ? 9 * size_t(K) : 0, 0.0);
    prio.assign(K, ~0ull);
    fill_until_stable();
    st = acvd::MeshState{};
    st.phase = acvd::kPhaseSweep;
    st.max_loops = std::min<int32_t>(max_loops, acvd::kMaxLoops - 2);
    st.max_convergences = max_conv;
    st.unconstrained_init = unconstrained_init;
    st.anchored = metric == 1 && !anchor.empty();
    st.first_rounds = first_rounds > 0 ? first_rounds : acvd::auto_first_rounds(N);
    st.constrained = unconstrained_init ? 0 : 1;
    st.early_stop_items = recompute(true);
    energy_log.clear();
    while (st.phase == acvd::kPhaseSweep) {
      pass();
      if (log_energy) energy_log.push_back(log_sum());
      if (!acvd::controller_step(st)) continue;
      int nd = clean();
      fill_until_stable();
      st.phase = acvd::finish_after_clean(st, nd) ? acvd::kPhaseDone : acvd::kPhaseSweep;
      recompute(true);
    }
    if (log_energy) final_energy = log_sum();
    stats.loops = st.loops;
    stats.convergences = st.convergences;
    stats.total_moves = st.total_mods;
    stats.last_disconnected = st.last_disconnected;
  }
};

}  // namespace oracle
