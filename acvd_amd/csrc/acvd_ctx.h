// Device-side data layout of one clustering context and the launchers of its
// kernels.  A context holds M >= 1 meshes packed back to back: items (mesh
// vertices) of mesh m occupy the global range [item_off, item_off + n_items)
// and its clusters [clus_off, clus_off + K).  Every kernel runs over the
// packed arrays, so a single mesh is simply the case M == 1.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "acvd_math.h"

namespace acvd {

constexpr int kBlock = 256;
// Move-counter slices: apply blocks add their count (no returned value) to
// counter blockIdx % kSlices, which spreads same-address atomics.
constexpr int kSlices = 64;
// Work-list slices: scan tile t appends to slice t % kWorkSlices, so that a
// slice counter sees few same-address atomics per pass.
constexpr int kWorkSlices = 2048;
// Upper bound of the block-owned segments (= evaluate / apply grid).
constexpr int kMaxSeg = 2048;
// Segments (evaluate / accept-round grid) of a large context: 4.5 blocks per
// CU of the 256.  The round kernels fit 3 to 8 blocks per CU (VGPRs and LDS),
// so a larger grid runs their dependent loads in a further wave of blocks.
// QEM torus of 8.39M vertices, K = 100k, to convergence: 2048 blocks
// 260.7 ms, 1536 245.0, 1280 243.8, 1024 245.4, 768 245.1, 512 261.1; with
// the later kernels 1408 241.2, 1280 241.2-241.6, 1152 239.5.
constexpr int kSweepSegs = 1152;
static_assert(kSweepSegs <= kMaxSeg, "segment index bound");
// Candidate lists of at most this many proposals (last pass) are walked one
// segment per block, longer ones evenly through an index of the segments.
constexpr int kSegWalkMax = 16384;
// Scan tiles: groups of kTileItems items of one mesh in spatial (Morton)
// order.  Each tile keeps a superset of the clusters its items and their
// neighbours belong to; a tile none of whose clusters changed in the last
// pass (and none of whose proposals lost) has no item to evaluate and is
// skipped after reading that set.
constexpr int kTileItems = 256;
constexpr int kTileSet = 64;  // set capacity; a larger set marks the tile always scanned
constexpr int kTilesPerScanBlock = kBlock / 64;  // k_scan: one wavefront per tile
// The sets are kept up to date (k_tile_update after each pass) only once a
// pass moves at most K / kTileOnDivisor items: before that nearly every tile
// holds a changed cluster, and every tile is marked always scanned instead
// (QEM torus of 8.39M vertices, K = 100k: K/32 263.0 ms, K/16 260.7 ms,
// K/8 260.9 ms to convergence).
constexpr int kTileOnDivisor = 16;

struct DevMesh {
  int64_t item_off;
  int64_t n_items;
  int64_t clus_off;
  int32_t K;
  int32_t metric;  // 0 isotropic, 1 QEM
  int32_t level;   // QEM quadrics optimisation level
  uint32_t seed;
  int32_t has_indicator;
  int32_t tile_off;  // first scan tile of the mesh
  double gradation;
  double clamp_ratio;
  Scales scales;
};

// Pass log entry: pass number, items evaluated, proposals, moves, moves of
// each accept round, and the phase flags of the first mesh (bit 0
// constrained, bit 1 connexity).
constexpr int kPassLogFields = 5 + kMatchRounds;
constexpr int kPassLogCap = 64;  // > passes per chunk

// One energy log line of a mesh: the energy after a pass and the device wall
// clock when it was summed.
struct EnergyLogEntry {
  double energy;
  unsigned long long clock;
};

// Small control block (device), read back by the host between pass chunks.
struct Ctrl {
  int32_t any_sweep;
  int32_t any_need_clean;
  int32_t fill_changed;
  int32_t seg_walk;  // candidate walk of the next pass: 1 segment per block, 0 dense
  int32_t mods[kMatchRounds][kSlices];  // applied moves of this pass, per accept round
  int32_t live[kMatchRounds][kSlices];  // candidates registered by the re-evaluate of round r
  int32_t stale[kMatchRounds][kSlices]; // candidates re-evaluated in round r
  int32_t recheck[kMatchRounds][kSlices]; // stale candidates whose own move was re-checked in round r
  unsigned long long evaluated;  // boundary items evaluated (all passes)
  unsigned long long proposed;   // candidates proposed (all passes)
  unsigned long long accepted;   // moves applied (all passes)
  unsigned long long accepted_constrained;  // of which in passes where mesh 0 was constrained (QEM solves)
  unsigned long long passes;     // passes run (any mesh sweeping)
  unsigned long long rounds;      // accept rounds that walked the candidates (all passes)
  unsigned long long walked;      // candidate entries read by the round kernels, dead ones included
  unsigned long long reeval_live; // live candidates registered by re-evaluate rounds
  unsigned long long reevaluated; // candidates re-evaluated (their clusters moved)
  unsigned long long rechecked;   // stale candidates whose own move was re-checked (the isotropic rule)
  // Per-pass counts of the last kPassLogCap passes (ring indexed by pass
  // number): the per-loop console output reads them between pass chunks.
  int32_t pass_log[kPassLogCap][kPassLogFields];
};

// Pointers handed to the sweep kernels (kernel argument block).
struct SweepArgs {
  int64_t n_items;
  int64_t n_clusters;
  int32_t n_meshes;
  int32_t n_tiles;
  const DevMesh* meshes;
  MeshState* state;
  const int32_t* item_mesh;  // null when n_meshes == 1
  const int32_t* clus_mesh;  // null when n_meshes == 1
  const int32_t* row_ptr;    // n_items + 1
  const int32_t* col;        // global item ids
  const unsigned long long* ring_adj;  // ring adjacency per item (ring_adjacency of acvd_math.h)
  const int32_t* anchor;     // per cluster: global id of its fixed item or -1; null when no mesh has any
  const int32_t* frozen;     // per cluster: nonzero when frozen (ForceManifold restarts); null when none is
  int32_t count_rounds;      // count the live and re-evaluated candidates of each round (timed runs)
  const float4* xw;          // position + weight
  const float* q;            // 10 floats per item (QEM meshes), may be null
  int32_t* labels;
  unsigned long long* acc_ws;  // 4 per cluster: W, Sx, Sy, Sz (int64 fixed)
  unsigned long long* acc_q;   // 10 per cluster (QEM), may be null
  double* evec;                // 9 per cluster (QEM): eigenvectors of the cluster quadric, may be null
  ClusterGeom* geom;           // centroid + W per cluster (FP64)
  double* T;                   // W |q - c|^2 per cluster (QEM), may be null
  int32_t* sizes;
  int32_t* stamp;              // mesh-local loop of the last modification
  unsigned long long* prio;    // accept keys per cluster
  int32_t* cand_pass;          // per item: loop of its last proposal
  // candidate records, indexed like cand:
  int32_t* cand_tgt;           // target of the live proposal (< 0: none)
  int32_t* cand_lab;           // label of the candidate item (fixed while it is live)
  int32_t* cand_since;         // loop * kMatchRounds + round of its last evaluation
  int32_t* cl_round;           // per cluster: loop * kMatchRounds + round of its last move
  int32_t* work;               // items to evaluate in this pass (kWorkSlices slices)
  int32_t* work_count;         // fill count of each work slice
  int32_t work_cap;            // capacity of one work slice
  Ctrl* ctrl;
  // Block-owned segments: evaluate block b appends the proposals of its
  // items to candidate segment b (entries [b * cand_cap, + seg_cand[b]));
  // apply block b appends the items it moved to moved segment b (capacity
  // cand_cap, per pass).
  int32_t n_seg;               // segments = evaluate / apply blocks
  int32_t cand_cap;
  int32_t* cand;               // item of each candidate
  int32_t* seg_cand;
  int32_t* moved;
  int32_t* seg_moved;
  // scan tiles
  const int32_t* tile_items;   // n_tiles * kTileItems item ids, -1 padding
  const int32_t* item_tile;    // tile of each item
  const int32_t* tile_mesh;    // mesh of each tile
  int32_t* tile_cnt;           // entries in the tile's cluster set (> kTileSet: overflow)
  int32_t* tile_set;           // n_tiles * kTileSet cluster ids
  int32_t* tile_prop;          // loop of the last proposal made in the tile
};

// ---- launchers (defined in the .hip files) --------------------------------
// build
hipError_t launch_csr_count(hipStream_t s, int64_t nf, const int32_t* tri_global, int32_t* deg2, int32_t* vfdeg);
hipError_t launch_csr_fill(hipStream_t s, int64_t nf, const int32_t* tri_global, const int32_t* tptr,
                           int32_t* tfill, int32_t* tcol, const int32_t* vfptr, int32_t* vffill, int32_t* vfcol);
hipError_t launch_csr_sort_unique(hipStream_t s, int64_t n, const int32_t* tptr, int32_t* tcol, int32_t* deg,
                                  const int32_t* vfptr, int32_t* vfcol);
hipError_t launch_csr_compact(hipStream_t s, int64_t n, const int32_t* tptr, const int32_t* tcol,
                              const int32_t* row_ptr, int32_t* col);
hipError_t exclusive_scan_i32(hipStream_t s, const int32_t* in, int32_t* out, int64_t n, void** tmp,
                              size_t* tmp_bytes);
hipError_t launch_ring_adj(hipStream_t s, int64_t n, const int32_t* row_ptr, const int32_t* col,
                            unsigned long long* ring_adj);
hipError_t launch_curvature(hipStream_t s, int64_t off, int64_t n, const double* pos, const int32_t* tri,
                            const int32_t* row_ptr, const int32_t* col, const int32_t* vfptr, const int32_t* vfcol,
                            double* out);
hipError_t launch_sample_topology(hipStream_t s, int64_t off, int64_t n, const double* pos, const int32_t* tri,
                                  const int32_t* vfptr, const int32_t* vfcol, const int32_t* row_ptr, int32_t* ring,
                                  int32_t* ring_len, int32_t* key, double* area);
hipError_t launch_item_build(hipStream_t s, const SweepArgs& a, const double* pos, const int32_t* tri_global,
                             const int32_t* vfptr, const int32_t* vfcol, const double* indicator,
                             double* raw_w, float4* xw, float* q);
hipError_t launch_seg_reduce(hipStream_t s, int n_blocks, const int2* block_tab, const SweepArgs& a,
                             const double* raw_w, int what, double* partial);
hipError_t launch_clamp_and_scales(hipStream_t s, const SweepArgs& a, int n_blocks, const int2* block_tab,
                                   const double* partial, int stage, const double* raw_w, DevMesh* meshes_rw);
// subdivision (one level: keys, sort, run heads, edge numbers and midpoints, faces)
hipError_t launch_sub_keys(hipStream_t s, int64_t nf, int64_t nv, const int32_t* tri, unsigned long long* keys,
                           int32_t* vals, int32_t* face_edge, int32_t* active);
hipError_t sort_sub_keys(hipStream_t s, int64_t n3, int64_t nv, unsigned long long* k0, unsigned long long* k1,
                         int32_t* v0, int32_t* v1, void** tmp, size_t* tmp_bytes);
hipError_t launch_sub_heads(hipStream_t s, int64_t n3, int64_t nv, const unsigned long long* keys,
                            const int32_t* vals, int32_t* flag);
hipError_t launch_sub_edges(hipStream_t s, int64_t n3, int64_t nv, const unsigned long long* keys,
                            const int32_t* vals, const int32_t* num, const int32_t* tri, const double* xyz,
                            double* xyz_out, const double* ind, double* ind_out, int32_t* face_edge);
hipError_t launch_sub_faces(hipStream_t s, int64_t nf, int64_t nv, const int32_t* tri, const int32_t* face_edge,
                            const int32_t* fpos, int32_t* tri_out);
// sweep
hipError_t launch_zero_clusters(hipStream_t s, const SweepArgs& a, const int32_t* mesh_mask);
hipError_t launch_accumulate(hipStream_t s, const SweepArgs& a, const int32_t* mesh_mask,
                             unsigned long long* mesh_count);
hipError_t launch_derive(hipStream_t s, const SweepArgs& a, const int32_t* mesh_mask, int set_stamp);
// tiles: the tile sets are kept (else every tile is marked always scanned).
hipError_t launch_scan(hipStream_t s, const SweepArgs& a, bool tiles);
hipError_t launch_tile_sets(hipStream_t s, const SweepArgs& a, const int32_t* mesh_mask);
hipError_t launch_tile_order(hipStream_t s, const SweepArgs& a, const float* bbox, int32_t* tile_items,
                             int32_t* item_tile, void** tmp, size_t* tmp_bytes);
hipError_t launch_eval(hipStream_t s, const SweepArgs& a, bool any_iso, bool any_qem, bool qem_unconstrained);
hipError_t launch_apply(hipStream_t s, const SweepArgs& a, int round, bool constrained);
hipError_t launch_tile_update(hipStream_t s, const SweepArgs& a);
hipError_t launch_reeval(hipStream_t s, const SweepArgs& a, bool any_iso, bool any_qem, bool qem_unconstrained,
                         int round);
hipError_t launch_controller(hipStream_t s, const SweepArgs& a);
hipError_t launch_fill_round(hipStream_t s, const SweepArgs& a, const int32_t* mesh_mask, int32_t* pending);
hipError_t launch_cc(hipStream_t s, const SweepArgs& a, const int32_t* mesh_mask, int32_t* parent,
                     int32_t* comp_size, unsigned long long* best, int32_t* ncomp, int32_t* ndisc);
hipError_t launch_seed_keys(hipStream_t s, const SweepArgs& a, int64_t off, int64_t n, uint32_t seed,
                            uint32_t* keys, int32_t* vals);
hipError_t launch_seed_assign(hipStream_t s, int32_t* labels, int64_t item_off, int64_t clus_off, int32_t K,
                              const int32_t* sorted_vals);
hipError_t sort_pairs_u32(hipStream_t s, const uint32_t* kin, uint32_t* kout, const int32_t* vin, int32_t* vout,
                          int64_t n, void** tmp, size_t* tmp_bytes);
hipError_t launch_cluster_points(hipStream_t s, const SweepArgs& a, int32_t mesh, double* pts);
hipError_t launch_item_energy(hipStream_t s, const SweepArgs& a, const double* pts_all, double* e_item);
// energy log: cluster energies into e_cluster (n_clusters), then per
// sweeping mesh entry `loops % ring` of its ring (ring entries per mesh), or
// with final_e every mesh's sum into final_e[mesh]
hipError_t launch_energy_log(hipStream_t s, const SweepArgs& a, double* e_cluster, EnergyLogEntry* log, int32_t ring,
                             double* final_e);
hipError_t launch_clock_stamp(hipStream_t s, unsigned long long* t);
hipError_t launch_fill_i32(hipStream_t s, int32_t* p, int64_t n, int32_t v);
// output mesh
hipError_t build_dual_faces(hipStream_t s, int64_t nf, const int32_t* tri, const int32_t* labels, int64_t item_off,
                            int64_t clus_off, int32_t* out_tri, int64_t* out_nf);
hipError_t fix_empty_cluster_points(hipStream_t s, const int32_t* sizes, int64_t clus_off, int32_t K, int32_t* first,
                                    double* pts);
hipError_t quadric_placement(hipStream_t s, int64_t n, int64_t item_off, const int32_t* labels, int64_t clus_off,
                             int32_t K, const float* q, double sq, unsigned long long* acc, int level, double* pts);
hipError_t launch_fill_u64(hipStream_t s, unsigned long long* p, int64_t n, unsigned long long v);

}  // namespace acvd
