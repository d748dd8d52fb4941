// Boundary sweep of the vertex clustering, one pass =
//   scan     : the boundary items whose clusters changed recently (or whose
//              last proposal lost) form the work list;
//   evaluate : each evaluates its best move to a neighbouring cluster
//              (isotropic or QEM energy, vtkUniformClustering.h:922-991) and,
//              if it lowers the energy, registers its accept key at the source
//              and target clusters (minimum per cluster) and joins the
//              candidate list;
//   rounds   : kMatchRounds accept rounds.  In each, a candidate whose key is
//              the minimum at both of its clusters is applied (label, int64
//              fixed-point accumulators of the two clusters, then their
//              centroid, W and FP64 rank-truncated QEM point); the remaining
//              candidates are re-evaluated against the new state when one of
//              their clusters moved, and register their next-round key;
//   controller: per-mesh phase logic of MinimizeEnergy
//              (vtkUniformClustering.h:724-830) after the pass.
// A cluster takes part in at most one applied move per round, so every
// applied dE is exact with respect to the state at the start of its round and
// the energy decreases monotonically.  Between convergences the host runs the
// clean (largest connected component per cluster, vtkUniformClustering.h:405-549)
// and fill (regrowth of null items, :551-633) kernels below.

#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include "acvd_ctx.h"

namespace acvd {

namespace {

inline unsigned grid_for(int64_t n) { return unsigned((n + kBlock - 1) / kBlock); }

__device__ inline int mesh_of_item(const SweepArgs& a, int64_t i) { return a.item_mesh ? a.item_mesh[i] : 0; }
__device__ inline int mesh_of_cluster(const SweepArgs& a, int64_t c) { return a.clus_mesh ? a.clus_mesh[c] : 0; }

__device__ inline int64_t ld_i64(const unsigned long long* p) { return int64_t(*p); }

// ----------------------------------------------------------------- statistics
__global__ void k_zero_clusters(SweepArgs a, const int32_t* __restrict__ mask) {
  int64_t c = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (c >= a.n_clusters) return;
  if (!mask[mesh_of_cluster(a, c)]) return;
  for (int k = 0; k < 4; ++k) a.acc_ws[4 * c + k] = 0ull;
  if (a.acc_q)
    for (int k = 0; k < 10; ++k) a.acc_q[10 * c + k] = 0ull;
  a.sizes[c] = 0;
}

// Item contributions in int64 fixed point (order-independent sums).  One
// block per scan tile (256 items in spatial order, a few clusters): the items
// first add into an LDS table keyed by cluster, then each table entry is
// flushed with one global atomic per value, instead of one per item.
static_assert(kBlock == kTileItems, "one thread per tile item");
__global__ __launch_bounds__(kBlock) void k_accumulate(SweepArgs a, const int32_t* __restrict__ mask,
                                                       unsigned long long* mesh_count) {
  constexpr int kVals = 15;  // W, S[3], Q[10], count
  __shared__ int32_t keys[kTileItems];
  __shared__ unsigned long long vals[kTileItems * kVals];
  __shared__ unsigned long long s_count;
  const int32_t t = blockIdx.x;
  const int m = a.item_mesh ? a.tile_mesh[t] : 0;
  if (!mask[m]) return;  // whole block
  keys[threadIdx.x] = -1;
  for (int k = 0; k < kVals; ++k) vals[threadIdx.x * kVals + k] = 0ull;
  if (threadIdx.x == 0) s_count = 0ull;
  __syncthreads();
  const DevMesh& dm = a.meshes[m];
  const int32_t i = a.tile_items[int64_t(t) * kTileItems + threadIdx.x];
  const int32_t c = i >= 0 ? a.labels[i] : kNull;
  if (c != kNull) {
    uint32_t h = (uint32_t(c) * 2654435761u) >> 24;
    int slot = 0;
    for (int probe = 0; probe < kTileItems; ++probe) {  // at most 256 keys: always finds a slot
      const int32_t old = atomicCAS(&keys[h], -1, c);
      if (old == -1 || old == c) {
        slot = int(h);
        break;
      }
      h = (h + 1) & (kTileItems - 1);
    }
    const float4 p = a.xw[i];
    const Scales sc = dm.scales;
    unsigned long long* v = vals + slot * kVals;
    atomicAdd(&v[0], (unsigned long long)fix_w(p.w, sc));
    atomicAdd(&v[1], (unsigned long long)fix_s(p.w, p.x, sc));
    atomicAdd(&v[2], (unsigned long long)fix_s(p.w, p.y, sc));
    atomicAdd(&v[3], (unsigned long long)fix_s(p.w, p.z, sc));
    if (dm.metric == 1)
      for (int k = 0; k < 10; ++k) atomicAdd(&v[4 + k], (unsigned long long)fix_q(a.q[10 * int64_t(i) + k], sc));
    atomicAdd(&v[14], 1ull);
    // early-stop items: those of clusters that are not frozen (:733-736)
    if (!(a.frozen && a.frozen[c])) atomicAdd(&s_count, 1ull);
  }
  __syncthreads();
  const int32_t key = keys[threadIdx.x];
  if (key >= 0) {
    const unsigned long long* v = vals + threadIdx.x * kVals;
    for (int k = 0; k < 4; ++k) atomicAdd(&a.acc_ws[4 * int64_t(key) + k], v[k]);
    if (dm.metric == 1)
      for (int k = 0; k < 10; ++k) atomicAdd(&a.acc_q[10 * int64_t(key) + k], v[4 + k]);
    atomicAdd(&a.sizes[key], int32_t(v[14]));
  }
  if (threadIdx.x == 0 && mesh_count && s_count) atomicAdd(&mesh_count[m], s_count);
}

// Centroid, W and (QEM, constrained) T of cluster c from its accumulators;
// the solve leaves the cluster's eigenvectors in evec, which warm-start the
// trial solves on the cluster and, with warm = true, this solve itself (the
// derive after a move, whose quadric differs from the last one by one item).
// kExtras = false: no mesh has fixed clusters (the anchor test is compiled out).
template <bool kExtras = true>
__device__ void derive_cluster(const SweepArgs& a, int64_t c, const DevMesh& dm, int constrained, bool warm) {
  const int64_t W = ld_i64(&a.acc_ws[4 * c]);
  const ClusterGeom g = geom_from_fixed(W, ld_i64(&a.acc_ws[4 * c + 1]), ld_i64(&a.acc_ws[4 * c + 2]),
                                        ld_i64(&a.acc_ws[4 * c + 3]), dm.scales);
  a.geom[c] = g;
  if (dm.metric == 1) {
    double T = 0.0;
    const int32_t an = (kExtras && a.anchor) ? a.anchor[c] : -1;
    if (an >= 0) {  // a fixed cluster: its point is its anchor item's position
      const float4 p = a.xw[an];
      T = anchored_T(g, AnchorPos{double(p.x), double(p.y), double(p.z), true});
    } else if (constrained) {
      int64_t q10[10];
      for (int k = 0; k < 10; ++k) q10[k] = ld_i64(&a.acc_q[10 * c + k]);
      double V0[9];
      if (warm)
        for (int k = 0; k < 9; ++k) V0[k] = a.evec[9 * c + k];
      T = qem_T_from_fixed(q10, g, dm.scales, dm.level, nullptr, warm ? V0 : nullptr, a.evec + 9 * c);
    }
    a.T[c] = T;
  }
}

__global__ void k_derive(SweepArgs a, const int32_t* __restrict__ mask, int set_stamp) {
  int64_t c = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (c >= a.n_clusters) return;
  const int m = mesh_of_cluster(a, c);
  if (!mask[m]) return;
  const DevMesh dm = a.meshes[m];
  derive_cluster(a, c, dm, a.state[m].constrained, false);
  if (set_stamp) a.stamp[c] = a.state[m].loops;  // SetAllClustersToModified
}

// ------------------------------------------------------------------- propose
// kExtras = false: no fixed or frozen cluster exists in the context, and the
// view says so at compile time (the kernels' anchor and freeze tests fold
// away; the host picks the instantiation).
template <bool kExtras>
__device__ inline SweepView make_view(const SweepArgs& a) {
  SweepView v;
  v.row_ptr = a.row_ptr;
  v.col = a.col;
  v.labels = a.labels;
  v.xw = reinterpret_cast<const float*>(a.xw);
  v.q = a.q;
  v.acc_ws = a.acc_ws;
  v.acc_q = a.acc_q;
  v.evec = a.evec;
  v.geom = a.geom;
  v.T = a.T;
  v.sizes = a.sizes;
  v.stamp = a.stamp;
  v.cand_pass = a.cand_pass;
  v.ring_adj = a.ring_adj;
  v.anchor = kExtras ? a.anchor : nullptr;
  v.frozen = kExtras ? a.frozen : nullptr;
  return v;
}

// needs_eval of acvd_math.h with the neighbour loads issued together (up to 8
// neighbours in registers) so that the label and stamp gathers overlap.
__device__ inline bool needs_eval_fast(const SweepArgs& a, int64_t i, int32_t lab, int32_t L) {
  const int32_t r0 = a.row_ptr[i], r1 = a.row_ptr[i + 1];
  bool dirty = (a.cand_pass[i] == L - 1);
  if (lab != kNull) dirty |= (a.stamp[lab] + 1 >= L);
  bool boundary = false;
  const int32_t deg = r1 - r0;
  int32_t nb[8], lb[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) nb[k] = k < deg ? a.col[r0 + k] : 0;
#pragma unroll
  for (int k = 0; k < 8; ++k) lb[k] = k < deg ? a.labels[nb[k]] : lab;
  // neighbour stamps are read only while the item is not already dirty
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    if (lb[k] != lab) {
      boundary = true;
      if (!dirty && lb[k] != kNull && a.stamp[lb[k]] + 1 >= L) dirty = true;
    }
  }
  for (int32_t k = r0 + 8; k < r1; ++k) {
    const int32_t b = a.labels[a.col[k]];
    if (b != lab) {
      boundary = true;
      if (!dirty && b != kNull && a.stamp[b] + 1 >= L) dirty = true;
    }
  }
  return boundary && dirty;
}

// Scan while the tile sets are off (nearly every tile holds a changed
// cluster): every item in id order, so that the item words and CSR rows are
// read in order.  Block b takes items [b * kBlock, (b + 1) * kBlock) and
// appends to work slice b % kWorkSlices (at most one tile's worth per block,
// within the slice capacity).  The evaluate's results do not depend on the
// order of the work list.  C3 torus: 139 -> 127 us per full scan.
__global__ __launch_bounds__(kBlock) void k_scan_linear(SweepArgs a) {
  __shared__ int32_t s_cnt[kBlock / 64], s_base;
  const int64_t i = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  bool want = false;
  if (i < a.n_items) {
    const MeshState& st = a.state[mesh_of_item(a, i)];
    if (st.phase == kPhaseSweep) want = needs_eval_fast(a, i, a.labels[i], st.loops);
  }
  const unsigned long long bal = __ballot(want);
  if (lane == 0) s_cnt[w] = __popcll(bal);
  __syncthreads();
  if (threadIdx.x == 0) {
    int tot = 0;
    for (int k = 0; k < kBlock / 64; ++k) {
      const int c = s_cnt[k];
      s_cnt[k] = tot;
      tot += c;
    }
    s_base = tot ? atomicAdd(&a.work_count[blockIdx.x % kWorkSlices], tot) : 0;
  }
  __syncthreads();
  if (want) {
    const int32_t pos = s_base + s_cnt[w] + __popcll(bal & ((1ull << lane) - 1ull));
    a.work[int64_t(blockIdx.x % kWorkSlices) * a.work_cap + pos] = int32_t(i);
  }
}

// Scan: items of sweeping meshes that must be evaluated this pass are
// compacted into the work list (slice t % kWorkSlices of tile t).  One
// wavefront per tile: the lanes first test the tile's cluster set against the
// cluster stamps (and the tile's last proposal loop); only an active tile
// reads its items and CSR rows.  The test is conservative: every item
// needs_eval would select lies in an active tile.
// It runs while the tile sets are kept; before that k_scan_linear scans every
// item.
static_assert(kTileSet == 64, "k_scan: one set entry per lane");
__global__ __launch_bounds__(kBlock) void k_scan(SweepArgs a) {
  // one wavefront per tile, kTilesPerScanBlock tiles per block (the waves
  // are independent: no block-level synchronisation)
  const int32_t t = int32_t(blockIdx.x) * kTilesPerScanBlock + int32_t(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (t >= a.n_tiles) return;  // wave-uniform
  const int m = a.item_mesh ? a.tile_mesh[t] : 0;
  // the tile's count, last proposal and set entry of this lane (the set holds
  // kTileSet = 64 entries; entries past cnt are ignored) read with the state
  const int32_t cnt = a.tile_cnt[t];
  const int32_t prop = a.tile_prop[t];
  const int32_t c_lane = a.tile_set[int64_t(t) * kTileSet + lane];
  const MeshState& st = a.state[m];
  if (st.phase != kPhaseSweep) return;
  const int32_t L = st.loops;
  bool active = cnt > kTileSet || prop == L - 1;
  if (!active) {  // wave-uniform branch
    bool hit = false;
    if (lane < cnt) hit = a.stamp[c_lane] + 1 >= L;
    active = __ballot(hit) != 0ull;
  }
  if (!active) return;
  constexpr int R = kTileItems / 64;
  int32_t it[R];
  bool want[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    it[r] = a.tile_items[int64_t(t) * kTileItems + r * 64 + lane];
    want[r] = it[r] >= 0 && needs_eval_fast(a, it[r], a.labels[it[r]], L);
  }
  unsigned long long bal[R];
  int32_t tot = 0;
#pragma unroll
  for (int r = 0; r < R; ++r) {
    bal[r] = __ballot(want[r]);
    tot += __popcll(bal[r]);
  }
  if (tot == 0) return;
  const int slice = t % kWorkSlices;
  int32_t base = 0;
  if (lane == 0) base = atomicAdd(&a.work_count[slice], tot);
  base = __shfl(base, 0);
  const unsigned long long below = (1ull << lane) - 1ull;
  int32_t off = 0;
#pragma unroll
  for (int r = 0; r < R; ++r) {
    if (want[r]) a.work[int64_t(slice) * a.work_cap + base + off + __popcll(bal[r] & below)] = it[r];
    off += __popcll(bal[r]);
  }
}

// Cluster set of a tile, rebuilt from scratch: labels of its items and of
// their neighbours, gathered in an LDS hash table (linear probing, bounded).
__global__ __launch_bounds__(64) void k_tile_sets(SweepArgs a, const int32_t* __restrict__ mask) {
  constexpr int kTab = 256;
  constexpr int32_t kEmpty = -2;
  __shared__ int32_t tab[kTab];
  __shared__ int32_t n_uniq, overflow, n_out;
  const int32_t t = blockIdx.x;
  const int m = a.item_mesh ? a.tile_mesh[t] : 0;
  if (mask && !mask[m]) return;
  for (int k = threadIdx.x; k < kTab; k += 64) tab[k] = kEmpty;
  if (threadIdx.x == 0) {
    n_uniq = 0;
    overflow = 0;
    n_out = 0;
  }
  __syncthreads();
  auto insert = [&](int32_t c) {
    if (c == kNull) return;
    uint32_t h = (uint32_t(c) * 2654435761u) >> 24;  // top 8 bits
    for (int probe = 0; probe < kTab; ++probe) {
      const int32_t old = atomicCAS(&tab[h], kEmpty, c);
      if (old == kEmpty) {
        atomicAdd(&n_uniq, 1);
        return;
      }
      if (old == c) return;
      h = (h + 1) & (kTab - 1);
    }
    atomicOr(&overflow, 1);
  };
  for (int r = 0; r < kTileItems / 64; ++r) {
    const int32_t i = a.tile_items[int64_t(t) * kTileItems + r * 64 + threadIdx.x];
    if (i < 0) continue;
    insert(a.labels[i]);
    for (int32_t k = a.row_ptr[i]; k < a.row_ptr[i + 1]; ++k) insert(a.labels[a.col[k]]);
  }
  __syncthreads();
  const bool big = overflow || n_uniq > kTileSet;
  if (!big) {
    for (int k = threadIdx.x; k < kTab; k += 64) {
      const int32_t c = tab[k];
      if (c != kEmpty) a.tile_set[int64_t(t) * kTileSet + atomicAdd(&n_out, 1)] = c;
    }
    // the free slots hold a value no cluster id equals: tile_insert may read a
    // slot another insert has reserved but not written yet, and must not
    // mistake a stale id left there for its own cluster
    for (int k = threadIdx.x; k < kTileSet; k += 64)
      if (k >= n_uniq) a.tile_set[int64_t(t) * kTileSet + k] = kEmpty;
  }
  if (threadIdx.x == 0) a.tile_cnt[t] = big ? kTileSet + 1 : n_uniq;
}
// ------------------------------------------------------------------ propose
// Work slices: a scan tile appends its work items to slice t % kWorkSlices;
// many slices keep the same-address counter atomics few.  A consumer block
// maps a position of the concatenation to its entry with an LDS index of the
// slice starts.
struct WorkIndex {
  int32_t start[kWorkSlices + 1];
};

__device__ inline void load_work_index(const int32_t* counts, WorkIndex* wi) {
  constexpr int kPer = kWorkSlices / kBlock;
  static_assert(kWorkSlices % kBlock == 0, "work slices per thread");
  __shared__ int32_t s_wave[kBlock / 64];
  int32_t v[kPer];
  int32_t tot = 0;
  for (int k = 0; k < kPer; ++k) {
    v[k] = counts[threadIdx.x * kPer + k];
    tot += v[k];
  }
  // inclusive scan of the thread totals over the wave, then over the waves
  const int lane = __lane_id(), wv = threadIdx.x >> 6;
  int32_t inc = tot;
  for (int off = 1; off < 64; off <<= 1) {
    const int32_t u = __shfl_up(inc, off);
    if (lane >= off) inc += u;
  }
  if (lane == 63) s_wave[wv] = inc;
  __syncthreads();
  int32_t before = inc - tot;
  for (int w = 0; w < wv; ++w) before += s_wave[w];
  for (int k = 0; k < kPer; ++k) {
    wi->start[threadIdx.x * kPer + k] = before;
    before += v[k];
  }
  if (threadIdx.x == kBlock - 1) wi->start[kWorkSlices] = before;
  __syncthreads();
}

__device__ inline int32_t work_entry(const SweepArgs& a, const WorkIndex* wi, int32_t p) {
  int lo = 0, hi = kWorkSlices;  // slice k with start[k] <= p < start[k+1]
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (wi->start[mid] <= p) lo = mid;
    else hi = mid;
  }
  return a.work[int64_t(lo) * a.work_cap + (p - wi->start[lo])];
}

// Candidate segments seen as one list: an LDS index of the segment starts
// (prefix of the n_seg <= kMaxSeg fill counts), so that consumer blocks walk
// all candidates evenly whatever their segments.
struct SegIndex {
  int32_t start[kMaxSeg + 1];
};

__device__ inline void load_seg_index(const int32_t* counts, int32_t n_seg, SegIndex* si) {
  constexpr int kPer = kMaxSeg / kBlock;
  __shared__ int32_t s_wave[kBlock / 64];
  int32_t v[kPer];
  int32_t tot = 0;
  for (int k = 0; k < kPer; ++k) {
    const int32_t j = threadIdx.x * kPer + k;
    v[k] = j < n_seg ? counts[j] : 0;
    tot += v[k];
  }
  const int lane = __lane_id(), wv = threadIdx.x >> 6;
  int32_t inc = tot;
  for (int off = 1; off < 64; off <<= 1) {
    const int32_t u = __shfl_up(inc, off);
    if (lane >= off) inc += u;
  }
  if (lane == 63) s_wave[wv] = inc;
  __syncthreads();
  int32_t before = inc - tot;
  for (int w = 0; w < wv; ++w) before += s_wave[w];
  for (int k = 0; k < kPer; ++k) {
    si->start[threadIdx.x * kPer + k] = before;
    before += v[k];
  }
  if (threadIdx.x == kBlock - 1) si->start[kMaxSeg] = before;
  __syncthreads();
}

// Storage index (segment * cap + offset) of position p of the list.
__device__ inline int64_t seg_pos(const SegIndex* si, int32_t cap, int32_t p) {
  int lo = 0, hi = kMaxSeg;  // segment k with start[k] <= p < start[k+1]
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (si->start[mid] <= p) lo = mid;
    else hi = mid;
  }
  return int64_t(lo) * cap + (p - si->start[lo]);
}

// True (in every thread of the block) when round r applied no move: the
// later rounds of the pass then apply nothing either and are skipped.
__device__ inline bool round_was_empty(const SweepArgs& a, int r) {
  __shared__ int s_empty;
  if (threadIdx.x < 64) {
    int m = a.ctrl->mods[r][threadIdx.x];
    for (int off = 32; off > 0; off >>= 1) m += __shfl_xor(m, off);
    if (threadIdx.x == 0) s_empty = m == 0;
  }
  __syncthreads();
  return s_empty != 0;
}

// True when round r lies past the pass's last round for the only mesh of the
// context (a batch checks each candidate's mesh instead).
__device__ inline bool past_last_round(const SweepArgs& a, int r) {
  return !a.item_mesh && r >= pass_rounds(a.state[0]);
}

// Metric of the meshes a kernel instantiation sweeps: 0 isotropic, 1 QEM,
// 2 QEM while unconstrained (isotropic rule, no FP64 solves compiled in).
constexpr int mesh_metric(int kernel_metric) { return kernel_metric == 2 ? 1 : kernel_metric; }

// Walks the candidate list, calling f(valid, entry) for one position per
// thread per step (every thread of the block takes part in every step).  A
// long list (last pass's proposals above kSegWalkMax) is walked evenly by all
// blocks through the LDS index; a short one by block b over segment b alone,
// which spares the index (grid = n_seg).  The results do not depend on the
// walk.
// Control values of the candidate walk, loaded by a kernel before its first
// barrier so that they share a memory round trip with its other first loads.
struct WalkCtl {
  bool by_seg;    // this pass walks one segment per block
  int32_t n_own;  // fill count of this block's segment
};
__device__ inline WalkCtl load_walk_ctl(const SweepArgs& a) {
  WalkCtl w;
  w.by_seg = a.ctrl->seg_walk != 0;
  w.n_own = a.seg_cand[blockIdx.x];
  return w;
}

template <class F>
__device__ inline void walk_candidates(const SweepArgs& a, const WalkCtl& w, SegIndex* si, F&& f) {
  int32_t n, first, stride;
  if (w.by_seg) {  // block-uniform
    n = w.n_own;
    first = 0;
    stride = kBlock;
  } else {
    load_seg_index(a.seg_cand, a.n_seg, si);
    n = si->start[kMaxSeg];
    first = blockIdx.x * kBlock;
    stride = gridDim.x * kBlock;
  }
  const int64_t seg = int64_t(blockIdx.x) * a.cand_cap;
  for (int32_t base = first; base < n; base += stride) {
    const int32_t k = base + int32_t(threadIdx.x);
    int64_t e = 0;
    if (k < n) e = w.by_seg ? seg + k : seg_pos(si, a.cand_cap, k);
    f(k < n, e);
  }
}

// Block-local minimum of the accept keys registered per cluster: an
// open-addressing table in LDS, flushed with one global atomicMin per distinct
// cluster.  Candidates of a block lie close together on the mesh and share
// few clusters, so this replaces most same-address global atomics.  The
// minimum does not depend on the order, so the keys are those of direct
// atomics.  A block registers at most 2 * kBlock keys between flushes, which
// keeps the table at most half full (every probe sequence ends).
constexpr int kKeyAggBits = 10;
constexpr int kKeyAggSlots = 1 << kKeyAggBits;
static_assert(2 * kBlock <= kKeyAggSlots / 2, "key table must stay at most half full");
struct KeyAgg {
  int32_t cl[kKeyAggSlots];
  unsigned long long key[kKeyAggSlots];
};

__device__ inline void keyagg_clear(KeyAgg* g) {
  for (int k = threadIdx.x; k < kKeyAggSlots; k += blockDim.x) {
    g->cl[k] = -1;
    g->key[k] = ~0ull;
  }
}

__device__ inline void keyagg_min(KeyAgg* g, int32_t c, unsigned long long key) {
  uint32_t h = (uint32_t(c) * 2654435761u) >> (32 - kKeyAggBits);
  while (true) {
    const int32_t prev = atomicCAS(&g->cl[h], -1, c);
    if (prev == -1 || prev == c) {
      atomicMin(&g->key[h], key);
      return;
    }
    h = (h + 1) & (kKeyAggSlots - 1);
  }
}

// Flushes the table to the global keys and clears it; all threads of the
// block call it (it synchronises the block before and after).
__device__ inline void keyagg_flush(KeyAgg* g, unsigned long long* prio) {
  __syncthreads();
  for (int k = threadIdx.x; k < kKeyAggSlots; k += blockDim.x) {
    const int32_t c = g->cl[k];
    if (c >= 0) {
      atomicMin(&prio[c], g->key[k]);
      g->cl[k] = -1;
      g->key[k] = ~0ull;
    }
  }
  __syncthreads();
}

// Wave-aggregated append of v to a block-owned list segment (fill count in
// LDS); every lane of the wave calls it.  Returns the entry index (when want).
__device__ inline int64_t seg_append(int64_t seg, int32_t* s_n, bool want) {
  const unsigned long long bal = __ballot(want);
  if (bal == 0ull) return -1;
  const int lane = __lane_id();
  const int leader = __ffsll((long long)bal) - 1;
  int32_t base = 0;
  if (lane == leader) base = atomicAdd(s_n, __popcll(bal));
  base = __shfl(base, leader);
  return want ? seg + base + __popcll(bal & ((1ull << lane) - 1ull)) : -1;
}

// ------------------------------------------------------ QEM solve pool (wave)
// The constrained QEM rule needs one FP64 solve for the item leaving its
// cluster and one per open target (qem_propose_short_row runs them one after
// the other in the item's lane).  Here the lanes of a wave put these solves
// into a shared pool and take them one per lane: a wave whose items need S
// solves in all runs ceil(S / 64) solve steps instead of as many as its
// busiest lane, and an item's solves run side by side.  The choice among the
// targets is the lexicographic minimum of (dE, cluster id), which does not
// depend on the order the solves finish in, so the move is the one of
// qem_propose_short_row.
constexpr int kPoolTasks = 64 * (1 + kRingCache);  // solves of one wave at most
struct QemPool {
  int32_t cl[kPoolTasks];     // target cluster, or ~a for the leave solve of cluster a
  uint8_t owner[kPoolTasks];  // lane of the item
  // results of one solve step (one per lane)
  float tb[64];   // join term of the target
  double Tn[64];  // T of the trial cluster
  double Tb[64];  // T of the target
};

// Orders the LDS accesses of the lanes of one wave (their DS operations
// execute in issue order; this keeps the compiler from moving them).
__device__ inline void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Wave-inclusive prefix sum of v.
__device__ inline int wave_prefix(int v) {
  const int lane = __lane_id();
  for (int off = 1; off < 64; off <<= 1) {
    const int u = __shfl_up(v, off);
    if (lane >= off) v += u;
  }
  return v;
}

// Owner kinds of the pool: a labelled item of a constrained QEM mesh whose
// screen passed, with a row of at most kRingCache neighbours (open targets in
// nb / open) or a longer one (open targets found again by walking the row).
enum PoolKind : int { kPoolNone = 0, kPoolShort = 1, kPoolLong = 2 };

// Screen of a labelled item of a constrained QEM mesh for the pool: its kind
// (kPoolNone when no target is open) and its number of solves.
__device__ inline int pool_screen(const SweepView& v, int64_t i, int32_t lab, int connexity, int32_t* nb,
                                  QemItem* s, uint32_t* open, int* ntask) {
  *ntask = 0;
  if (v.row_ptr[i + 1] - v.row_ptr[i] <= kRingCache) {
    if (!qem_screen_short_row(v, i, lab, connexity, 1, nb, s, open)) return kPoolNone;
    *ntask = 1 + __popc(*open);
    return kPoolShort;
  }
  if (!qem_screen(v, i, lab, connexity, 1, s)) return kPoolNone;
  int n = 1;
  for (int32_t k = qem_next_open(v, lab, *s, 1, s->r0 - 1); k < s->r1; k = qem_next_open(v, lab, *s, 1, k)) ++n;
  *ntask = n;
  return kPoolLong;
}

// Posts solves [j0, j1) of a long-row owner at pool index first (solve 0 is
// the leave, solve j > 0 the j-th open target in CSR order).
__device__ inline void pool_post_long(const SweepView& v, QemPool* pool, int64_t i, int32_t lab, int j0, int j1,
                                      int first) {
  QemItem s;
  qem_item_terms(v, i, lab, 1, &s);
  int32_t k = s.r0 - 1;
  for (int j = 1; j < j0; ++j) k = qem_next_open(v, lab, s, 1, k);
  for (int j = j0; j < j1; ++j) {
    int32_t c = ~lab;
    if (j > 0) {
      k = qem_next_open(v, lab, s, 1, k);
      c = k < s.r1 ? v.labels[v.col[k]] : ~lab;  // the state is frozen: the count of pool_screen holds
    }
    pool->cl[first + j - j0] = c;
    pool->owner[first + j - j0] = uint8_t(__lane_id());
  }
}

// Every lane of the wave calls it.  An owner (kind != kPoolNone, from
// pool_screen, mesh m) receives its item's move as qem_propose_short_row /
// the generic QEM rule would return it.  Short-row owners post all their
// solves first (a wave posts at most kPoolTasks of them), so that the lanes
// keep only the choice state across the solve steps; long-row owners post
// after them what fits and the rest in later fills.
template <bool kExtras>
__device__ bool qem_pool_choose(const SweepArgs& a, const SweepView& v, QemPool* pool, int kind, int ntask,
                                int64_t i, int32_t lab, int32_t m, const int32_t* nb, const QemItem& s,
                                uint32_t open, int32_t* target_out) {
  const int lane = __lane_id();
  int first, total, post;
  {
    const int cs = kind == kPoolShort ? ntask : 0, cl = kind == kPoolLong ? ntask : 0;
    const int is = wave_prefix(cs), il = wave_prefix(cl);
    const int ts = __shfl(is, 63), tl = __shfl(il, 63);
    first = kind == kPoolShort ? is - cs : ts + il - cl;
    total = ts + tl < kPoolTasks ? ts + tl : kPoolTasks;
    post = total - first < ntask ? total - first : ntask;
    if (post < 0) post = 0;
  }
  if (kind == kPoolShort) {
    pool->cl[first] = ~lab;
    pool->owner[first] = uint8_t(lane);
    for (int j = 1; j < ntask; ++j) {
      pool->cl[first + j] = ring_label(nb, __builtin_ctz(open));
      pool->owner[first + j] = uint8_t(lane);
      open &= open - 1;
    }
  } else if (kind == kPoolLong && post > 0) {
    pool_post_long(v, pool, i, lab, 0, post, first);
  }
  const float pw = s.pw, ta = s.ta;
  const double Ta = s.Ta;
  double Ta_new = 0.0, best_d = 0.0, best_m = 0.0;
  int32_t target = kNull;
  int done = 0;
  while (total > 0) {
    wave_lds_sync();
    for (int base = 0; base < total; base += 64) {
      const bool solve = lane < total - base;
      const int ow = solve ? pool->owner[base + lane] : lane;
      const int64_t ti = __shfl(int32_t(i), ow);
      const int32_t tm = __shfl(m, ow);
      if (solve) {
        const int32_t tc = pool->cl[base + lane];
        const int32_t c = tc < 0 ? ~tc : tc;
        const DevMesh& dm = a.meshes[tm];
        const Scales sc = dm.scales;
        const AnchorPos an = kExtras ? anchor_pos(v, c) : AnchorPos{0.0, 0.0, 0.0, false};
        double Tn = 0.0;  // a centroid point (an unconstrained mesh's unanchored cluster)
        if (an.valid || a.state[tm].constrained) {
          int64_t iw, is[3], iq[10], ws[4], q10[10];
          item_fixed(v, ti, sc, &iw, is, iq);
          load_acc(v, c, ws, q10);
          Tn = qem_T_trial(ws, q10, iw, is, iq, tc < 0 ? -1 : 1, sc, dm.level, v.evec + 9 * int64_t(c), an);
        }
        pool->Tn[lane] = Tn;
        if (tc >= 0) {
          pool->tb[lane] =
              join_term(v.xw[4 * ti], v.xw[4 * ti + 1], v.xw[4 * ti + 2], v.xw[4 * ti + 3], v.geom[tc]);
          pool->Tb[lane] = v.T[tc];
        }
      }
      wave_lds_sync();
      // collect this lane's solves among pool entries [base, base + 64)
      const int j0 = first > base ? first : base;
      const int j1 = first + post < base + 64 ? first + post : base + 64;
      for (int j = j0; j < j1; ++j) {
        if (j - first + done == 0) {
          Ta_new = pool->Tn[j - base];
          continue;
        }
        const int32_t b = pool->cl[j];
        const float tb = pool->tb[j - base];
        const double Tb_new = pool->Tn[j - base], Tb = pool->Tb[j - base];
        const double d = qem_delta(pw, ta, tb, Ta_new, Ta, Tb_new, Tb);
        if (target == kNull || d < best_d || (d == best_d && b < target)) {
          best_d = d;
          best_m = qem_mag(pw, ta, tb, Ta_new, Ta, Tb_new, Tb);
          target = b;
        }
      }
      wave_lds_sync();
    }
    // next fill: long-row solves that did not fit
    done += post;
    const int rem = kind == kPoolLong ? ntask - done : 0;
    const int inc = wave_prefix(rem);
    const int tot = __shfl(inc, 63);
    first = inc - rem;
    total = tot < kPoolTasks ? tot : kPoolTasks;
    post = total - first < rem ? total - first : rem;
    if (post < 0) post = 0;
    if (post > 0) pool_post_long(v, pool, i, lab, done, done + post, first);
  }
  *target_out = target;
  return target != kNull && qem_accept(best_d, best_m);
}

// Evaluate (round 0): best move of each work item (shared rule of
// acvd_math.h) against the state at the pass start; a proposal registers its
// round-0 accept key at both clusters and joins the block's candidate
// segment.  Block b owns segment b (capacity cand_cap, count seg_cand[b]);
// the second launch of a mixed batch (append = 1) continues the segments of
// the first.
template <int METRIC, bool kExtras>
__global__ __launch_bounds__(kBlock) void k_eval(SweepArgs a, int append) {
  __shared__ WorkIndex wi;
  __shared__ KeyAgg agg;
  __shared__ QemPool pools[METRIC == 1 ? kBlock / 64 : 1];
  __shared__ int32_t s_n;
  if (threadIdx.x == 0) s_n = append ? a.seg_cand[blockIdx.x] : 0;
  keyagg_clear(&agg);
  load_work_index(a.work_count, &wi);
  const int32_t n = wi.start[kWorkSlices];
  const SweepView v = make_view<kExtras>(a);
  const int64_t seg = int64_t(blockIdx.x) * a.cand_cap;
  for (int32_t base = blockIdx.x * kBlock; base < n; base += gridDim.x * kBlock) {
    const int32_t t = base + int32_t(threadIdx.x);
    bool want = false;
    int32_t i = 0, lab = kNull, target = kNull, L = 0, m = 0;
    int32_t nb[kRingCache];
    QemItem s;
    uint32_t open = 0;
    int kind = kPoolNone, ntask = 0;
    if (t < n) {
      i = work_entry(a, &wi, t);
      m = mesh_of_item(a, i);
      const DevMesh& dm = a.meshes[m];
      if (dm.metric == mesh_metric(METRIC)) {
        const MeshState* st = &a.state[m];
        L = st->loops;
        lab = a.labels[i];
        if (METRIC == 1 && lab != kNull && qem_mode(*st))
          kind = pool_screen(v, i, lab, st->connexity, nb, &s, &open, &ntask);
        else  // no FP64 solve: the null-item rule or the isotropic one
          want = propose_item<METRIC == 1 ? 2 : METRIC>(v, i, lab, st->connexity, st->constrained, dm.level,
                                                        dm.scales, &target);
      }
    }
    if constexpr (METRIC == 1) {
      int32_t pt = kNull;
      const bool r =
          qem_pool_choose<kExtras>(a, v, &pools[threadIdx.x >> 6], kind, ntask, i, lab, m, nb, s, open, &pt);
      if (kind != kPoolNone) {
        want = r;
        target = pt;
      }
    }
    if (want) {
      const DevMesh& dm = a.meshes[m];
      const unsigned long long key = accept_key(dm.seed, L, 0, uint32_t(i - dm.item_off));
      if (lab != kNull) keyagg_min(&agg, lab, key);
      keyagg_min(&agg, target, key);
      a.cand_pass[i] = L;
      a.tile_prop[a.item_tile[i]] = L;
      if (a.item_mesh) atomicAdd(&a.state[m].pass_props, 1);
    }
    const int64_t e = seg_append(seg, &s_n, want);
    if (want) {
      a.cand[e] = i;
      a.cand_lab[e] = lab;
      a.cand_tgt[e] = target;
      a.cand_since[e] = L * kMatchRounds;
    }
    keyagg_flush(&agg, a.prio);
  }
  __syncthreads();
  if (threadIdx.x == 0) a.seg_cand[blockIdx.x] = s_n;
}

// Re-evaluation of candidate entry e (when valid) for round `round`; every
// lane of the wave calls it (the QEM solve pool is shared by the wave).
// reg(c, key) registers an accept key at cluster c.
// Block sum of v (the result is valid in thread 0; synchronises the block).
__device__ inline int block_sum(int v) {
  __shared__ int s_red[kBlock / 64];
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  if (__lane_id() == 0) s_red[threadIdx.x / 64] = v;
  __syncthreads();
  int tot = 0;
  if (threadIdx.x == 0)
    for (int w = 0; w < kBlock / 64; ++w) tot += s_red[w];
  __syncthreads();
  return tot;  // valid in thread 0
}

template <int METRIC, bool kExtras, bool kCount, class Reg>
__device__ inline void reeval_entry(const SweepArgs& a, const SweepView& v, QemPool* pool, bool valid, int64_t e,
                                    int round, Reg&& reg, int* n_live, int* n_stale, int* n_recheck) {
  // the candidate record in one round trip
  const int32_t b = valid ? a.cand_tgt[e] : -1;
  const int32_t i_e = valid ? a.cand[e] : 0;
  const int32_t lab_e = valid ? a.cand_lab[e] : kNull;
  const int32_t since = valid ? a.cand_since[e] : 0;
  bool live = false, stale = false, use_pool = false, recheck = false;
  int32_t i = 0, lab = kNull, target = b, m = 0;
  unsigned long long key = 0;
  int32_t nb[kRingCache];
  QemItem s;
  uint32_t open = 0;
  int kind = kPoolNone, ntask = 0;
  if (b >= 0) {
    i = i_e;
    lab = lab_e;  // a live candidate has not moved
    m = mesh_of_item(a, i);
    const DevMesh& dm = a.meshes[m];
    const MeshState& st = a.state[m];
    // a mesh past its last round of the pass registers nothing (no move); a
    // context of one mesh has left the kernel already (past_last_round)
    if (dm.metric == mesh_metric(METRIC) && (!a.item_mesh || round < pass_rounds(st))) {
      const int32_t ra = lab != kNull ? a.cl_round[lab] : -1, rb = a.cl_round[b];
      stale = ra >= since || rb >= since;
      key = accept_key(dm.seed, st.loops, round, uint32_t(i - dm.item_off));
      live = true;
      if (stale) {
        a.cand_since[e] = st.loops * kMatchRounds + round;
        target = kNull;
        use_pool = METRIC == 1 && lab != kNull && qem_mode(st);
        if (use_pool)
          kind = pool_screen(v, i, lab, st.connexity, nb, &s, &open, &ntask);
        else if (lab != kNull) {  // the isotropic rule: the move itself is checked again
          target = b;
          recheck = true;
          live = iso_recheck(v, i, lab, b, st.connexity);
        } else  // the null-item rule
          live = propose_item<METRIC == 1 ? 2 : METRIC>(v, i, lab, st.connexity, st.constrained, dm.level,
                                                        dm.scales, &target);
      }
    }
  }
  if constexpr (METRIC == 1) {
    int32_t pt = kNull;
    const bool r = qem_pool_choose<kExtras>(a, v, pool, kind, ntask, i, lab, m, nb, s, open, &pt);
    if (use_pool) {
      live = r;
      target = pt;
    }
  }
  if (stale) a.cand_tgt[e] = live ? target : -3;
  if (live) {
    if (lab != kNull) reg(lab, key);
    reg(target, key);
  }
  if constexpr (kCount) {
    *n_live += live;
    *n_stale += stale && !recheck;
    *n_recheck += recheck;
  }
}

// Re-evaluate (round r > 0): a live candidate (cand_tgt >= 0) whose own
// cluster or target moved since its last evaluation is evaluated again
// against the state left by round r-1: it registers its round-r key with its
// (possibly new) target or is dropped (-3).  The others keep their move, whose
// dE depends only on the two unchanged clusters, and register it again.
// kCount: count the registered and re-evaluated candidates (timed runs).
template <int METRIC, bool kExtras, bool kCount>
__global__ __launch_bounds__(kBlock) void k_reeval(SweepArgs a, int round) {
  __shared__ KeyAgg agg;
  __shared__ SegIndex si;
  __shared__ QemPool pools[METRIC == 1 ? kBlock / 64 : 1];
  const WalkCtl w = load_walk_ctl(a);
  const bool past = past_last_round(a, round);
  keyagg_clear(&agg);  // made visible by the barrier in round_was_empty
  if (round_was_empty(a, round - 1) || past) return;
  const SweepView v = make_view<kExtras>(a);
  QemPool* pool = &pools[METRIC == 1 ? threadIdx.x >> 6 : 0];
  int n_live = 0, n_stale = 0, n_recheck = 0;
  walk_candidates(a, w, &si, [&](bool valid, int64_t e) {
    reeval_entry<METRIC, kExtras, kCount>(
        a, v, pool, valid, e, round, [&](int32_t c, unsigned long long key) { keyagg_min(&agg, c, key); },
        &n_live, &n_stale, &n_recheck);
    keyagg_flush(&agg, a.prio);
  });
  // candidates registered and re-evaluated in this round (statistics of the
  // timed runs: the block sums cost a few barriers per round)
  if constexpr (kCount) {
    const int tl = block_sum(n_live);
    const int ts = block_sum(n_stale);
    const int tr = block_sum(n_recheck);
    if (threadIdx.x == 0) {
      if (tl) atomicAdd(&a.ctrl->live[round][blockIdx.x % kSlices], tl);
      if (ts) atomicAdd(&a.ctrl->stale[round][blockIdx.x % kSlices], ts);
      if (tr) atomicAdd(&a.ctrl->recheck[round][blockIdx.x % kSlices], tr);
    }
  }
}

// --------------------------------------------------------------------- apply
// Adds cluster c to the set of tile t unless present (sets only grow between
// rebuilds; duplicates from concurrent inserts only use capacity).
__device__ inline void tile_insert(const SweepArgs& a, int32_t t, int32_t c) {
  // the count and all set entries read at once (16-byte loads; the set holds
  // kTileSet entries), entries past cnt ignored
  const int32_t cnt = a.tile_cnt[t];
  const int4* set4 = reinterpret_cast<const int4*>(a.tile_set + int64_t(t) * kTileSet);
  int4 sv[kTileSet / 4];
#pragma unroll
  for (int q = 0; q < kTileSet / 4; ++q) sv[q] = set4[q];
  if (cnt > kTileSet) return;
  bool found = false;
#pragma unroll
  for (int q = 0; q < kTileSet / 4; ++q) {
    const int4 v = sv[q];
    found |= (4 * q < cnt && v.x == c) || (4 * q + 1 < cnt && v.y == c) || (4 * q + 2 < cnt && v.z == c) ||
             (4 * q + 3 < cnt && v.w == c);
  }
  if (found) return;
  const int32_t slot = atomicAdd(&a.tile_cnt[t], 1);
  if (slot < kTileSet) a.tile_set[int64_t(t) * kTileSet + slot] = c;
}


// Apply round `round`: a live candidate holding the minimum round key at both
// of its clusters moves: label, fixed-point accumulator deltas (plain
// read-modify-write: the round owns both clusters), sizes and stamps; then the
// centroid, W and (QEM) T of both clusters are derived from their final
// accumulators (the round's only writer of them).  The blocks walk all
// candidates evenly; the item joins its block's moved segment (tile sets are
// extended by k_tile_update after the pass).
//  kSolve = false (no mesh constrained): a derive is a centroid, computed by
//    the mover itself;
//  kSolve = true: a derive may need an FP64 solve; the clusters the movers of
//    a wave modified are listed in LDS and derived one per lane.
struct DeriveList {
  int32_t cl[2 * 64];
};

// Apply of candidate entry e (when valid) in round `round`; every lane of the
// wave calls it.  Returns true when the candidate moved (item *moved_item).
template <bool kSolve, bool kExtras>
__device__ inline bool apply_entry(const SweepArgs& a, DeriveList* dl, bool valid, int64_t e, int round,
                                   int32_t* moved_item) {
  bool win = false;
  int64_t i = 0;
  int32_t lab = kNull;
  // the candidate record in one round trip
  const int32_t b = valid ? a.cand_tgt[e] : -1;
  const int32_t i_e = valid ? a.cand[e] : 0;
  const int32_t lab_e = valid ? a.cand_lab[e] : kNull;
  if (b >= 0) {
    i = i_e;
    lab = lab_e;
    const int m = mesh_of_item(a, i);
    MeshState* st = &a.state[m];
    const int32_t L = st->loops;
    const DevMesh& dm = a.meshes[m];
    const unsigned long long key = accept_key(dm.seed, L, round, uint32_t(i - dm.item_off));
    const unsigned long long pa = lab != kNull ? a.prio[lab] : key, pb = a.prio[b];
    win = pa == key && pb == key;
    if (win) {
      // exclusive owner of clusters lab and b for this round
      a.labels[i] = b;
      a.cand_tgt[e] = -3;
      const float4 p = a.xw[i];
      const Scales sc = dm.scales;
      const unsigned long long iw = (unsigned long long)fix_w(p.w, sc);
      const unsigned long long is[3] = {(unsigned long long)fix_s(p.w, p.x, sc),
                                        (unsigned long long)fix_s(p.w, p.y, sc),
                                        (unsigned long long)fix_s(p.w, p.z, sc)};
      const int32_t rtag = L * kMatchRounds + round;
      if (lab != kNull) {
        unsigned long long* wa = a.acc_ws + 4 * int64_t(lab);
        wa[0] -= iw;
        for (int q = 0; q < 3; ++q) wa[1 + q] -= is[q];
        a.sizes[lab] -= 1;
        a.stamp[lab] = L;
        a.cl_round[lab] = rtag;
      }
      unsigned long long* wb = a.acc_ws + 4 * int64_t(b);
      wb[0] += iw;
      for (int q = 0; q < 3; ++q) wb[1 + q] += is[q];
      // quadric accumulators: while no mesh is constrained (kSolve = false)
      // nothing reads them, and the recompute at the convergence that
      // constrains a mesh rebuilds them from the labels
      if (kSolve && dm.metric == 1) {
        unsigned long long* qa = lab != kNull ? a.acc_q + 10 * int64_t(lab) : nullptr;
        unsigned long long* qb = a.acc_q + 10 * int64_t(b);
        for (int q = 0; q < 10; ++q) {
          const unsigned long long iq = (unsigned long long)fix_q(a.q[10 * i + q], sc);
          if (qa) qa[q] -= iq;
          qb[q] += iq;
        }
      }
      a.sizes[b] += 1;
      a.stamp[b] = L;
      a.cl_round[b] = rtag;
      if (a.item_mesh) atomicAdd(&st->pass_mods, 1);
      if (!kSolve) {
        if (lab != kNull) derive_cluster<kExtras>(a, lab, dm, 0, false);
        derive_cluster<kExtras>(a, b, dm, 0, false);
      }
    }
  }
  if constexpr (kSolve) {
    const int cnt = win ? (lab != kNull ? 2 : 1) : 0;
    const int inc = wave_prefix(cnt);
    const int total = __shfl(inc, 63);
    if (cnt) {
      dl->cl[inc - cnt] = b;
      if (cnt == 2) dl->cl[inc - 1] = lab;
    }
    // the movers' accumulator stores before the other lanes' loads (one
    // CU: workgroup scope orders them through its L1)
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    for (int k = __lane_id(); k < total; k += 64) {
      const int64_t c = dl->cl[k];
      const int m = mesh_of_cluster(a, c);
      derive_cluster<kExtras>(a, c, a.meshes[m], a.state[m].constrained, true);
    }
    wave_lds_sync();
  }
  *moved_item = int32_t(i);
  return win;
}

template <bool kSolve, bool kExtras>
__global__ __launch_bounds__(kBlock) void k_apply(SweepArgs a, int round) {
  __shared__ int32_t s_moved;
  __shared__ SegIndex si;
  __shared__ DeriveList dlist[kSolve ? kBlock / 64 : 1];
  const int64_t mseg = int64_t(blockIdx.x) * a.cand_cap;
  const WalkCtl w = load_walk_ctl(a);
  const int32_t moved0 = round == 0 ? 0 : a.seg_moved[blockIdx.x];
  const bool past = past_last_round(a, round);
  if (round > 0 && (round_was_empty(a, round - 1) || past)) return;
  if (threadIdx.x == 0) s_moved = moved0;
  __syncthreads();
  DeriveList* dl = &dlist[kSolve ? threadIdx.x >> 6 : 0];
  int mods = 0;
  walk_candidates(a, w, &si, [&](bool valid, int64_t e) {
    int32_t i;
    const bool win = apply_entry<kSolve, kExtras>(a, dl, valid, e, round, &i);
    mods += win;
    const int64_t em = seg_append(mseg, &s_moved, win);
    if (em >= 0) a.moved[em] = i;
  });
  const int tm = block_sum(mods);  // synchronises the block
  if (threadIdx.x == 0) {
    a.seg_moved[blockIdx.x] = s_moved;
    if (tm) atomicAdd(&a.ctrl->mods[round][blockIdx.x % kSlices], tm);
  }
}

// Tile sets after a pass: an item that moved to cluster b shows b to its own
// tile and to the tiles of its neighbours (k_scan of the next pass reads them).
__global__ __launch_bounds__(kBlock) void k_tile_update(SweepArgs a) {
  const int32_t n = a.seg_moved[blockIdx.x];
  const int64_t seg = int64_t(blockIdx.x) * a.cand_cap;
  for (int32_t k = threadIdx.x; k < n; k += kBlock) {
    const int64_t i = a.moved[seg + k];
    const int32_t b = a.labels[i];
    const int32_t ti = a.item_tile[i];
    const int32_t r0 = a.row_ptr[i], r1 = a.row_ptr[i + 1];
    // the tiles of the first kRingCache neighbours read at once; each distinct
    // tile other than the item's gets one insert
    int32_t tj[kRingCache];
    ACVD_UNROLL
    for (int p = 0; p < kRingCache; ++p) tj[p] = r0 + p < r1 ? a.item_tile[a.col[r0 + p]] : ti;
    tile_insert(a, ti, b);
    ACVD_UNROLL
    for (int p = 0; p < kRingCache; ++p) {
      bool first = tj[p] != ti;
      ACVD_UNROLL
      for (int q = 0; q < p; ++q) first = first && tj[q] != tj[p];
      if (first) tile_insert(a, tj[p], b);
    }
    for (int32_t r = r0 + kRingCache; r < r1; ++r) {
      const int32_t t = a.item_tile[a.col[r]];
      if (t != ti) tile_insert(a, t, b);
    }
  }
}

// ---------------------------------------------------------------- controller
__global__ void k_controller(SweepArgs a) {
  __shared__ int any_sweep, any_clean;
  if (threadIdx.x == 0) {
    any_sweep = 0;
    any_clean = 0;
  }
  __syncthreads();
  __shared__ int s_work, s_props, s_mods, s_round_mods[kMatchRounds];
  __shared__ int s_part[2][kBlock / 64];
  static_assert(kSlices == 64, "one wave sums the slice counters");
  {  // items evaluated and proposals of the pass: sums over the work slices
     // (cleared for the next scan) and the candidate segments
    int w = 0, p = 0;
    for (int k = threadIdx.x; k < kWorkSlices; k += blockDim.x) {
      w += a.work_count[k];
      a.work_count[k] = 0;
    }
    for (int k = threadIdx.x; k < a.n_seg; k += blockDim.x) p += a.seg_cand[k];
    for (int off = 32; off > 0; off >>= 1) {
      w += __shfl_xor(w, off);
      p += __shfl_xor(p, off);
    }
    if (__lane_id() == 0) {
      s_part[0][threadIdx.x >> 6] = w;
      s_part[1][threadIdx.x >> 6] = p;
    }
  }
  __shared__ int s_live, s_stale, s_recheck;
  if (threadIdx.x < 64) {  // lane k reads slice k of the move counters, wave sums
    int dr[kMatchRounds];
    int lv = 0, st = 0, rc = 0;
    for (int r = 0; r < kMatchRounds; ++r) {
      dr[r] = a.ctrl->mods[r][threadIdx.x];
      lv += a.ctrl->live[r][threadIdx.x];
      st += a.ctrl->stale[r][threadIdx.x];
      rc += a.ctrl->recheck[r][threadIdx.x];
    }
    for (int off = 32; off > 0; off >>= 1) {
      for (int r = 0; r < kMatchRounds; ++r) dr[r] += __shfl_xor(dr[r], off);
      lv += __shfl_xor(lv, off);
      st += __shfl_xor(st, off);
      rc += __shfl_xor(rc, off);
    }
    if (threadIdx.x == 0) {
      int d = 0;
      for (int r = 0; r < kMatchRounds; ++r) {
        s_round_mods[r] = dr[r];
        d += dr[r];
      }
      s_mods = d;
      s_live = lv;
      s_stale = st;
      s_recheck = rc;
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    int w = 0, p = 0;
    for (int k = 0; k < kBlock / 64; ++k) {
      w += s_part[0][k];
      p += s_part[1][k];
    }
    s_work = w;
    s_props = p;
  }
  __syncthreads();
  bool ran = false;
  __shared__ int s_flags;  // phase flags of mesh 0 during this pass
  __shared__ int s_rounds; // most accept rounds a sweeping mesh ran this pass
  if (threadIdx.x == 0) s_rounds = 1;
  __syncthreads();
  for (int m = threadIdx.x; m < a.n_meshes; m += blockDim.x) {
    MeshState s = a.state[m];
    if (m == 0) s_flags = (s.constrained ? 1 : 0) | (s.connexity ? 2 : 0);
    if (s.phase == kPhaseSweep) {
      atomicMax(&s_rounds, pass_rounds(s));
      ran = true;
      if (!a.item_mesh) {  // one mesh: the sliced counters are its own
        s.pass_props = s_props;
        s.pass_mods = s_mods;
      }
      controller_step(s);
      a.state[m] = s;
    }
    if (s.phase == kPhaseSweep) atomicOr(&any_sweep, 1);
    if (s.phase == kPhaseNeedClean) atomicOr(&any_clean, 1);
  }
  if (ran) atomicOr(&any_sweep, 2);
  __syncthreads();
  if (threadIdx.x < kSlices)
    for (int r = 0; r < kMatchRounds; ++r) {
      a.ctrl->mods[r][threadIdx.x] = 0;
      a.ctrl->live[r][threadIdx.x] = 0;
      a.ctrl->stale[r][threadIdx.x] = 0;
      a.ctrl->recheck[r][threadIdx.x] = 0;
    }
  if (threadIdx.x == 0 && (any_sweep & 2)) {
    const unsigned long long pass = a.ctrl->passes;
    int32_t* e = a.ctrl->pass_log[pass % kPassLogCap];
    e[0] = int32_t(pass);
    e[1] = s_work;
    e[2] = s_props;
    e[3] = s_mods;
    for (int r = 0; r < kMatchRounds; ++r) e[4 + r] = s_round_mods[r];
    e[4 + kMatchRounds] = s_flags;
  }
  if (threadIdx.x == 0) {
    a.ctrl->evaluated += (unsigned long long)s_work;
    a.ctrl->proposed += (unsigned long long)s_props;
    a.ctrl->accepted += (unsigned long long)s_mods;
    if (s_flags & 1) a.ctrl->accepted_constrained += (unsigned long long)s_mods;
    // accept rounds that walked the candidate list: round 0, and each later
    // round whose predecessor applied a move (a round after one without moves
    // exits at once); the re-evaluate kernels walked the list in all but one
    if (any_sweep & 2) {
      int walked = 1;
      while (walked < s_rounds && s_round_mods[walked - 1] > 0) ++walked;
      a.ctrl->rounds += (unsigned long long)walked;
      a.ctrl->walked += (unsigned long long)s_props * (unsigned long long)(2 * walked - 1);
      a.ctrl->reeval_live += (unsigned long long)s_live;
      a.ctrl->reevaluated += (unsigned long long)s_stale;
      a.ctrl->rechecked += (unsigned long long)s_recheck;
    }
    a.ctrl->any_sweep = any_sweep & 1;
    a.ctrl->seg_walk = s_props <= kSegWalkMax ? 1 : 0;
    a.ctrl->any_need_clean = any_clean;
    if (any_sweep & 2) a.ctrl->passes += 1;
  }
}

// ---------------------------------------------------------------------- fill
// One synchronous round: every null item with a labelled neighbour takes the
// label of its lowest-id labelled neighbour (as of the start of the round).
__global__ void k_fill_pick(SweepArgs a, const int32_t* __restrict__ mask, int32_t* pending) {
  int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  bool changed = false;
  if (i < a.n_items && mask[mesh_of_item(a, i)] && a.labels[i] == kNull) {
    int32_t pick = kNull;
    for (int32_t k = a.row_ptr[i]; k < a.row_ptr[i + 1]; ++k) {
      const int32_t b = a.labels[a.col[k]];
      if (b != kNull) {
        pick = b;
        break;
      }
    }
    pending[i] = pick;
    changed = pick != kNull;
  }
  if (__ballot(changed) != 0 && __lane_id() == 0) a.ctrl->fill_changed = 1;
}

__global__ void k_fill_commit(SweepArgs a, const int32_t* __restrict__ mask, const int32_t* __restrict__ pending) {
  int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= a.n_items || !mask[mesh_of_item(a, i)] || a.labels[i] != kNull) return;
  a.labels[i] = pending[i];
}

// --------------------------------------------------- connected components
// Union-find over same-label edges; a root is always the smallest item id of
// its tree (larger roots are hooked under smaller ones), so the final
// representative of a component is its minimum item id.
__device__ int32_t cc_find(int32_t* parent, int32_t x) {
  while (true) {
    int32_t p = parent[x];
    if (p == x) return x;
    int32_t gp = parent[p];
    if (gp != p) parent[x] = gp;  // path halving (benign race: ancestors only)
    x = gp;
  }
}

__global__ void k_cc_init(SweepArgs a, const int32_t* __restrict__ mask, int32_t* parent, int32_t* comp_size) {
  int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= a.n_items || !mask[mesh_of_item(a, i)]) return;
  parent[i] = int32_t(i);
  comp_size[i] = 0;
}

__global__ void k_cc_hook(SweepArgs a, const int32_t* __restrict__ mask, int32_t* parent) {
  int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= a.n_items || !mask[mesh_of_item(a, i)]) return;
  const int32_t lab = a.labels[i];
  if (lab == kNull) return;
  for (int32_t k = a.row_ptr[i]; k < a.row_ptr[i + 1]; ++k) {
    const int32_t j = a.col[k];
    if (j <= i || a.labels[j] != lab) continue;
    int32_t u = int32_t(i), v = j;
    while (true) {
      u = cc_find(parent, u);
      v = cc_find(parent, v);
      if (u == v) break;
      if (u > v) {
        int32_t t = u;
        u = v;
        v = t;
      }
      // hook root v under the smaller root u
      if (atomicCAS(&parent[v], v, u) == v) break;
    }
  }
}

__global__ void k_cc_count(SweepArgs a, const int32_t* __restrict__ mask, int32_t* parent, int32_t* comp_size) {
  int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= a.n_items || !mask[mesh_of_item(a, i)]) return;
  if (a.labels[i] == kNull) return;
  const int32_t r = cc_find(parent, int32_t(i));
  atomicAdd(&comp_size[r], 1);
}

__global__ void k_cc_best(SweepArgs a, const int32_t* __restrict__ mask, const int32_t* __restrict__ parent,
                          const int32_t* __restrict__ comp_size, unsigned long long* best, int32_t* ncomp) {
  int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= a.n_items || !mask[mesh_of_item(a, i)]) return;
  const int32_t lab = a.labels[i];
  if (lab == kNull || parent[i] != i) return;
  // the component holding the cluster's fixed item first (its size counts 1e9
  // in CleanClustering, vtkUniformClustering.h:441-447), then the larger
  // component, then the smaller first item
  uint32_t size = uint32_t(comp_size[i]);
  const int32_t an = a.anchor ? a.anchor[lab] : -1;
  if (an >= 0) {
    int32_t r = an;
    while (parent[r] != r) r = parent[r];  // read-only find (the hooks are done)
    if (r == int32_t(i)) size |= 0x80000000u;
  }
  const unsigned long long key = (uint64_t(size) << 32) | uint64_t(0xFFFFFFFFu - uint32_t(i));
  atomicMax(&best[lab], key);
  atomicAdd(&ncomp[lab], 1);
}

__global__ void k_cc_reset(SweepArgs a, const int32_t* __restrict__ mask, int32_t* parent,
                           const unsigned long long* __restrict__ best) {
  int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= a.n_items || !mask[mesh_of_item(a, i)]) return;
  const int32_t lab = a.labels[i];
  if (lab == kNull) return;
  const int32_t r = cc_find(parent, int32_t(i));
  const uint32_t keep = 0xFFFFFFFFu - uint32_t(best[lab] & 0xFFFFFFFFull);
  if (uint32_t(r) != keep) a.labels[i] = kNull;
}

__global__ void k_cc_zero(SweepArgs a, const int32_t* __restrict__ mask, unsigned long long* best, int32_t* ncomp) {
  int64_t c = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (c >= a.n_clusters || !mask[mesh_of_cluster(a, c)]) return;
  best[c] = 0ull;
  ncomp[c] = 0;
}

__global__ void k_cc_disc(SweepArgs a, const int32_t* __restrict__ mask, const int32_t* __restrict__ ncomp,
                          int32_t* ndisc) {
  int64_t c = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (c >= a.n_clusters) return;
  const int m = mesh_of_cluster(a, c);
  if (!mask[m]) return;
  if (ncomp[c] > 1) atomicAdd(&ndisc[m], 1);
}

// ----------------------------------------------------------------- seeding
// Weighted random seeds: key_i = E_i / w_i with E_i ~ Exp(1); the K smallest
// keys of a mesh (one sort) seed its clusters, so that denser regions get
// more seeds.
__global__ void k_seed_keys(SweepArgs a, int64_t off, int64_t n, uint32_t seed, uint32_t* keys, int32_t* vals) {
  int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int64_t i = off + t;
  uint32_t h = prio_hash(seed ^ 0x5EEDu, 0x9E37u, uint32_t(t));
  double u = (double(h) + 0.5) * (1.0 / 4294967296.0);
  double w = a.xw[i].w;
  float k = float(-log(u) / (w > 0 ? w : 1e-300));
  keys[t] = __float_as_uint(k);  // positive floats order like their bits
  vals[t] = int32_t(t);
}

__global__ void k_seed_assign(int32_t* labels, int64_t item_off, int64_t clus_off, int32_t K,
                              const int32_t* __restrict__ sorted_vals) {
  int64_t r = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (r >= K) return;
  labels[item_off + sorted_vals[r]] = int32_t(clus_off + r);
}

// ----------------------------------------------------------------- outputs
// Cluster point of mesh `mesh`: centroid (isotropic) or the QEM point when the
// metric is constrained (GetClusterCentroid semantics).
__global__ void k_cluster_points(SweepArgs a, int32_t mesh, double* pts) {
  const DevMesh dm = a.meshes[mesh];
  int64_t r = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (r >= dm.K) return;
  const int64_t c = dm.clus_off + r;
  const int64_t W = ld_i64(&a.acc_ws[4 * c]);
  const ClusterGeom g = geom_from_fixed(W, ld_i64(&a.acc_ws[4 * c + 1]), ld_i64(&a.acc_ws[4 * c + 2]),
                                        ld_i64(&a.acc_ws[4 * c + 3]), dm.scales);
  double q[3] = {g.cx, g.cy, g.cz};
  const int32_t an = a.anchor && dm.metric == 1 ? a.anchor[c] : -1;
  if (!(g.W > 0.0)) {
    q[0] = q[1] = q[2] = __builtin_nan("");
  } else if (an >= 0) {  // fixed cluster: the anchor (exact input coordinates are set by the host)
    const float4 p = a.xw[an];
    q[0] = p.x;
    q[1] = p.y;
    q[2] = p.z;
  } else if (dm.metric == 1 && a.state[mesh].constrained) {
    int64_t q10[10];
    for (int k = 0; k < 10; ++k) q10[k] = ld_i64(&a.acc_q[10 * c + k]);
    qem_T_from_fixed(q10, g, dm.scales, dm.level, q);
  }
  pts[3 * r] = q[0];
  pts[3 * r + 1] = q[1];
  pts[3 * r + 2] = q[2];
}

// Per-item unshifted energy w |x - p_c|^2 with p_c given per global cluster.
__global__ void k_item_energy(SweepArgs a, const double* __restrict__ pts_all, double* e_item) {
  int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= a.n_items) return;
  const int32_t c = a.labels[i];
  if (c == kNull) {
    e_item[i] = 0.0;
    return;
  }
  const float4 p = a.xw[i];
  double dx = double(p.x) - pts_all[3 * int64_t(c)];
  double dy = double(p.y) - pts_all[3 * int64_t(c) + 1];
  double dz = double(p.z) - pts_all[3 * int64_t(c) + 2];
  e_item[i] = double(p.w) * (dx * dx + dy * dy + dz * dz);
}

// Energy log, step 1 (WriteToGlobalEnergyLog, vtkUniformClustering.h:1383-1395):
// the energy of every cluster from its exact accumulators, the point solved
// from scratch (no warm start, so the value depends on the labels only).
__global__ void k_log_cluster_energy(SweepArgs a, double* __restrict__ e) {
  int64_t c = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (c >= a.n_clusters) return;
  const int m = mesh_of_cluster(a, c);
  const DevMesh dm = a.meshes[m];
  const int64_t W = ld_i64(&a.acc_ws[4 * c]);
  const ClusterGeom g = geom_from_fixed(W, ld_i64(&a.acc_ws[4 * c + 1]), ld_i64(&a.acc_ws[4 * c + 2]),
                                        ld_i64(&a.acc_ws[4 * c + 3]), dm.scales);
  double T = 0.0;
  if (dm.metric == 1) {
    const int32_t an = a.anchor ? a.anchor[c] : -1;
    if (an >= 0) {
      const float4 p = a.xw[an];
      T = anchored_T(g, AnchorPos{double(p.x), double(p.y), double(p.z), true});
    } else if (a.state[m].constrained) {
      int64_t q10[10];
      for (int k = 0; k < 10; ++k) q10[k] = ld_i64(&a.acc_q[10 * c + k]);
      T = qem_T_from_fixed(q10, g, dm.scales, dm.level, nullptr);
    }
  }
  e[c] = log_cluster_energy(g, T);
}

// Energy log, step 2: one block per mesh sums its clusters in the order of
// kLogThreads (acvd_math.h).  In a pass (final_e null) a sweeping mesh writes
// its energy and the device wall clock to entry `loops % ring` of its ring of
// log entries, before the controller counts the pass; the final call writes
// every mesh's energy to final_e.
__global__ void __launch_bounds__(kLogThreads) k_log_reduce(SweepArgs a, const double* __restrict__ e,
                                                            EnergyLogEntry* log, int32_t ring, double* final_e) {
  const int m = blockIdx.x;
  const int32_t phase = a.state[m].phase, loop = a.state[m].loops;
  if (!final_e && (phase != kPhaseSweep || loop < 0)) return;  // uniform over the block
  const DevMesh dm = a.meshes[m];
  __shared__ double part[kLogThreads];
  double s = 0.0;
  for (int64_t r = threadIdx.x; r < dm.K; r += kLogThreads) s += e[dm.clus_off + r];
  part[threadIdx.x] = s;
  __syncthreads();
  for (int h = kLogThreads / 2; h > 0; h >>= 1) {
    if (int(threadIdx.x) < h) part[threadIdx.x] += part[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    if (final_e) {
      final_e[m] = part[0];
    } else {
      EnergyLogEntry& x = log[int64_t(m) * ring + loop % ring];
      x.energy = part[0];
      x.clock = (unsigned long long)wall_clock64();
    }
  }
}

__global__ void k_clock_stamp(unsigned long long* t) { *t = (unsigned long long)wall_clock64(); }

__global__ void k_fill_i32(int32_t* p, int64_t n, int32_t v) {
  int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}
__global__ void k_fill_u64(unsigned long long* p, int64_t n, unsigned long long v) {
  int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

}  // namespace

// ------------------------------------------------------------------ launchers
hipError_t launch_zero_clusters(hipStream_t s, const SweepArgs& a, const int32_t* mask) {
  if (a.n_clusters == 0) return hipSuccess;
  hipLaunchKernelGGL(k_zero_clusters, dim3(grid_for(a.n_clusters)), dim3(kBlock), 0, s, a, mask);
  return hipGetLastError();
}

hipError_t launch_accumulate(hipStream_t s, const SweepArgs& a, const int32_t* mask, unsigned long long* mesh_count) {
  if (a.n_items == 0) return hipSuccess;
  hipLaunchKernelGGL(k_accumulate, dim3(unsigned(a.n_tiles)), dim3(kBlock), 0, s, a, mask, mesh_count);
  return hipGetLastError();
}

hipError_t launch_derive(hipStream_t s, const SweepArgs& a, const int32_t* mask, int set_stamp) {
  if (a.n_clusters == 0) return hipSuccess;
  hipLaunchKernelGGL(k_derive, dim3(grid_for(a.n_clusters)), dim3(kBlock), 0, s, a, mask, set_stamp);
  return hipGetLastError();
}

hipError_t launch_tile_sets(hipStream_t s, const SweepArgs& a, const int32_t* mesh_mask) {
  if (a.n_tiles == 0) return hipSuccess;
  hipLaunchKernelGGL(k_tile_sets, dim3(unsigned(a.n_tiles)), dim3(64), 0, s, a, mesh_mask);
  return hipGetLastError();
}

hipError_t launch_scan(hipStream_t s, const SweepArgs& a, bool tiles) {
  if (a.n_items == 0) return hipSuccess;
  const dim3 grid(unsigned((a.n_tiles + kTilesPerScanBlock - 1) / kTilesPerScanBlock));
  // blocks of kBlock items never outnumber the scan tiles (each mesh's tiles
  // cover its items), so the linear scan stays within the slice capacity
  static_assert(kBlock == kTileItems, "one scan block per tile's worth of items");
  if (tiles) hipLaunchKernelGGL(k_scan, grid, dim3(kBlock), 0, s, a);
  else hipLaunchKernelGGL(k_scan_linear, dim3(unsigned((a.n_items + kBlock - 1) / kBlock)), dim3(kBlock), 0, s, a);
  return hipGetLastError();
}

// One evaluate block per candidate segment: block b writes segment b's count.
// qem_unconstrained: no QEM mesh is constrained (the QEM meshes take the
// isotropic rule, METRIC 2).
// The kernels come in two builds: with the anchor and freeze tests (some mesh
// has fixed or frozen clusters) and without them.
template <bool kX>
void launch_eval_x(hipStream_t s, const SweepArgs& a, bool any_iso, bool any_qem, bool qem_unconstrained) {
  if (any_iso) hipLaunchKernelGGL((k_eval<0, kX>), dim3(unsigned(a.n_seg)), dim3(kBlock), 0, s, a, 0);
  if (any_qem && qem_unconstrained)
    hipLaunchKernelGGL((k_eval<2, kX>), dim3(unsigned(a.n_seg)), dim3(kBlock), 0, s, a, any_iso ? 1 : 0);
  else if (any_qem)
    hipLaunchKernelGGL((k_eval<1, kX>), dim3(unsigned(a.n_seg)), dim3(kBlock), 0, s, a, any_iso ? 1 : 0);
}

hipError_t launch_eval(hipStream_t s, const SweepArgs& a, bool any_iso, bool any_qem, bool qem_unconstrained) {
  if (a.anchor || a.frozen) launch_eval_x<true>(s, a, any_iso, any_qem, qem_unconstrained);
  else launch_eval_x<false>(s, a, any_iso, any_qem, qem_unconstrained);
  return hipGetLastError();
}

// One apply block per candidate segment (its moved segment has the same
// index).  constrained: some mesh is constrained (derives may need a solve).
hipError_t launch_apply(hipStream_t s, const SweepArgs& a, int round, bool constrained) {
  const dim3 g(unsigned(a.n_seg)), b(kBlock);
  const bool x = a.anchor || a.frozen;
  if (constrained && x) hipLaunchKernelGGL((k_apply<true, true>), g, b, 0, s, a, round);
  else if (constrained) hipLaunchKernelGGL((k_apply<true, false>), g, b, 0, s, a, round);
  else if (x) hipLaunchKernelGGL((k_apply<false, true>), g, b, 0, s, a, round);
  else hipLaunchKernelGGL((k_apply<false, false>), g, b, 0, s, a, round);
  return hipGetLastError();
}

hipError_t launch_tile_update(hipStream_t s, const SweepArgs& a) {
  hipLaunchKernelGGL(k_tile_update, dim3(unsigned(a.n_seg)), dim3(kBlock), 0, s, a);
  return hipGetLastError();
}

template <bool kX, bool kC>
void launch_reeval_x(hipStream_t s, const SweepArgs& a, bool any_iso, bool any_qem, bool qem_unconstrained,
                     int round) {
  const dim3 g(unsigned(a.n_seg)), b(kBlock);
  if (any_iso) hipLaunchKernelGGL((k_reeval<0, kX, kC>), g, b, 0, s, a, round);
  if (any_qem && qem_unconstrained) hipLaunchKernelGGL((k_reeval<2, kX, kC>), g, b, 0, s, a, round);
  else if (any_qem) hipLaunchKernelGGL((k_reeval<1, kX, kC>), g, b, 0, s, a, round);
}

hipError_t launch_reeval(hipStream_t s, const SweepArgs& a, bool any_iso, bool any_qem, bool qem_unconstrained,
                         int round) {
  const bool x = a.anchor || a.frozen, c = a.count_rounds != 0;
  if (x && c) launch_reeval_x<true, true>(s, a, any_iso, any_qem, qem_unconstrained, round);
  else if (x) launch_reeval_x<true, false>(s, a, any_iso, any_qem, qem_unconstrained, round);
  else if (c) launch_reeval_x<false, true>(s, a, any_iso, any_qem, qem_unconstrained, round);
  else launch_reeval_x<false, false>(s, a, any_iso, any_qem, qem_unconstrained, round);
  return hipGetLastError();
}

hipError_t launch_controller(hipStream_t s, const SweepArgs& a) {
  hipLaunchKernelGGL(k_controller, dim3(1), dim3(kBlock), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_fill_round(hipStream_t s, const SweepArgs& a, const int32_t* mask, int32_t* pending) {
  if (a.n_items == 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill_pick, dim3(grid_for(a.n_items)), dim3(kBlock), 0, s, a, mask, pending);
  hipLaunchKernelGGL(k_fill_commit, dim3(grid_for(a.n_items)), dim3(kBlock), 0, s, a, mask, pending);
  return hipGetLastError();
}

hipError_t launch_cc(hipStream_t s, const SweepArgs& a, const int32_t* mask, int32_t* parent, int32_t* comp_size,
                     unsigned long long* best, int32_t* ncomp, int32_t* ndisc) {
  if (a.n_items == 0) return hipSuccess;
  const dim3 gi(grid_for(a.n_items)), gc(grid_for(a.n_clusters)), b(kBlock);
  hipLaunchKernelGGL(k_cc_init, gi, b, 0, s, a, mask, parent, comp_size);
  hipLaunchKernelGGL(k_cc_zero, gc, b, 0, s, a, mask, best, ncomp);
  hipLaunchKernelGGL(k_cc_hook, gi, b, 0, s, a, mask, parent);
  hipLaunchKernelGGL(k_cc_count, gi, b, 0, s, a, mask, parent, comp_size);
  hipLaunchKernelGGL(k_cc_best, gi, b, 0, s, a, mask, parent, comp_size, best, ncomp);
  hipLaunchKernelGGL(k_cc_reset, gi, b, 0, s, a, mask, parent, best);
  hipLaunchKernelGGL(k_cc_disc, gc, b, 0, s, a, mask, ncomp, ndisc);
  return hipGetLastError();
}

hipError_t launch_seed_keys(hipStream_t s, const SweepArgs& a, int64_t off, int64_t n, uint32_t seed,
                            uint32_t* keys, int32_t* vals) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_seed_keys, dim3(grid_for(n)), dim3(kBlock), 0, s, a, off, n, seed, keys, vals);
  return hipGetLastError();
}

hipError_t launch_seed_assign(hipStream_t s, int32_t* labels, int64_t item_off, int64_t clus_off, int32_t K,
                              const int32_t* sorted_vals) {
  if (K == 0) return hipSuccess;
  hipLaunchKernelGGL(k_seed_assign, dim3(grid_for(K)), dim3(kBlock), 0, s, labels, item_off, clus_off, K,
                     sorted_vals);
  return hipGetLastError();
}

hipError_t sort_pairs_u32(hipStream_t s, const uint32_t* kin, uint32_t* kout, const int32_t* vin, int32_t* vout,
                          int64_t n, void** tmp, size_t* tmp_bytes) {
  size_t need = 0;
  hipError_t e = hipcub::DeviceRadixSort::SortPairs(nullptr, need, kin, kout, vin, vout, int(n), 0, 32, s);
  if (e != hipSuccess) return e;
  if (need > *tmp_bytes) {
    if (*tmp) (void)hipFree(*tmp);
    *tmp = nullptr;
    e = hipMalloc(tmp, need);
    if (e != hipSuccess) return e;
    *tmp_bytes = need;
  }
  return hipcub::DeviceRadixSort::SortPairs(*tmp, need, kin, kout, vin, vout, int(n), 0, 32, s);
}

hipError_t launch_cluster_points(hipStream_t s, const SweepArgs& a, int32_t mesh, double* pts) {
  hipLaunchKernelGGL(k_cluster_points, dim3(unsigned((a.n_clusters + kBlock - 1) / kBlock) + 1), dim3(kBlock), 0, s,
                     a, mesh, pts);
  return hipGetLastError();
}

hipError_t launch_item_energy(hipStream_t s, const SweepArgs& a, const double* pts_all, double* e_item) {
  if (a.n_items == 0) return hipSuccess;
  hipLaunchKernelGGL(k_item_energy, dim3(grid_for(a.n_items)), dim3(kBlock), 0, s, a, pts_all, e_item);
  return hipGetLastError();
}

hipError_t launch_energy_log(hipStream_t s, const SweepArgs& a, double* e_cluster, EnergyLogEntry* log, int32_t ring,
                             double* final_e) {
  if (a.n_clusters == 0 || a.n_meshes == 0) return hipSuccess;
  hipLaunchKernelGGL(k_log_cluster_energy, dim3(grid_for(a.n_clusters)), dim3(kBlock), 0, s, a, e_cluster);
  hipLaunchKernelGGL(k_log_reduce, dim3(unsigned(a.n_meshes)), dim3(kLogThreads), 0, s, a, e_cluster, log, ring,
                     final_e);
  return hipGetLastError();
}

hipError_t launch_clock_stamp(hipStream_t s, unsigned long long* t) {
  hipLaunchKernelGGL(k_clock_stamp, dim3(1), dim3(1), 0, s, t);
  return hipGetLastError();
}

hipError_t launch_fill_i32(hipStream_t s, int32_t* p, int64_t n, int32_t v) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill_i32, dim3(grid_for(n)), dim3(kBlock), 0, s, p, n, v);
  return hipGetLastError();
}

hipError_t launch_fill_u64(hipStream_t s, unsigned long long* p, int64_t n, unsigned long long v) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill_u64, dim3(grid_for(n)), dim3(kBlock), 0, s, p, n, v);
  return hipGetLastError();
}

}  // namespace acvd
