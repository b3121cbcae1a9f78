// Device-resident octree of one point set (uniform depth, non-empty cells only)
// and the interaction lists between a target tree and a source tree.
//
// Definitions (restated by the CPU oracle for a bit-exact comparison):
//   cell coords at level L: ix = clamp(floor((x - lo) * inv_cell), 0, 2^L - 1)
//   Morton code: bit b of ix -> bit 3b+2, of iy -> 3b+1, of iz -> 3b
//   points sorted stably by leaf code; cells of level l = unique(code >> 3(L-l))
//   P2P slot (dx+1)*9+(dy+1)*3+(dz+1), separation criterion 1
//   M2L: target t, source s at the same level with parents adjacent (or equal)
//        and s, t not adjacent; transfer vector v = (dx+3)*49+(dy+3)*7+(dz+3)
#pragma once

#include <vector>

#include "common.hpp"

namespace pam {

struct Box {
  double lo[3];
  double width;
};

struct TreeLevel {
  int64_t n = 0;
  DevBuf<uint64_t> codes;        // [n] sorted unique cell codes
  DevBuf<int32_t> parent;        // [n] parent index in level-1 (level >= 1)
  DevBuf<int32_t> child_start;   // [n+1] child ranges in level+1 (level < leaf level)
};

constexpr int kMaxTreeLevel = 20;

struct Tree {
  int level = -1;                // leaf level L
  int sort_level = -1;           // level of the Morton codes the points are sorted by (>= L)
  Box box{};
  double inv_cell = 0.0;         // 2^L / width
  int64_t n_points = 0;
  DevBuf<uint64_t> point_codes;  // [n] sorted leaf codes of the points
  DevBuf<int32_t> perm;          // [n] sorted position -> input index
  DevBuf<int32_t> point_leaf;    // [n] sorted position -> leaf index
  DevBuf<int32_t> leaf_start;    // [n_leaves+1]
  DevBuf<int64_t> open_hist;     // [sort_level+1] points whose coarsest newly opened cell is at level l
  std::vector<TreeLevel> levels; // [0..L]; levels[L] are the leaves (buffers of deeper entries kept for reuse)
  int64_t n_leaves() const { return level >= 0 ? levels[level].n : 0; }
};

// Scratch buffers reused across builds.
struct TreeWorkspace {
  DevBuf<int32_t> flags, scan, scan_scratch, hist, tv;
  DevBuf<uint64_t> tk, tmp_codes;
  DevBuf<int32_t> tmp_idx;
  DevBuf<int64_t> counts;  // per-level cell counts while a tree is built
  DevBuf<int64_t> layout;  // per-level M2L column layouts while lists are built
};

struct M2LLevel {
  int64_t n_cols = 0;             // target columns, padded to kM2LCols per octant group
  int64_t n_cells = 0;            // target cells of the level
  DevBuf<int32_t> col_cell;       // [n_cols] target cell or -1
  DevBuf<int32_t> blk_oct;        // [n_cols/kM2LCols] octant of each column block
  std::vector<int32_t> h_blk_oct; // host copy of blk_oct
  DevBuf<int32_t> src;            // [n_trg_cells][189] source cell or -1 (octant slot order)
  // Pair-major form: the (target cell, slot) pairs grouped by transfer vector
  // v (0..342), each v's group padded to whole blocks of kM2LCols pairs. The
  // order of the pairs inside a group is immaterial: every pair's transfer is
  // computed on its own and the per-target sum runs in slot order.
  int64_t n_blk = 0;                // blocks of all groups
  std::vector<int32_t> h_v_blocks;  // [343] blocks of each v
  std::vector<int32_t> h_v_first;   // [343] first block of each v
  DevBuf<int32_t> blk_src;          // [n_blk * kM2LCols] source cell of each pair, -1 for padding
  DevBuf<int32_t> pair_pos;         // [n_trg_cells][189] pair index of (cell, slot), -1 if none
};

struct Lists {
  int level = -1;                 // leaf level the lists were built for (-1: none)
  DevBuf<int32_t> p2p;            // [n_trg_leaves][27]
  std::vector<M2LLevel> m2l;      // index = level (entries < 2 unused; may hold more levels than used)
  DevBuf<int32_t> v_count;        // [level + 1][343] pairs per transfer vector (then fill counters)
  DevBuf<int32_t> v_first;        // [level + 1][343] first block of each transfer vector
  HostStage<int32_t> v_stage;     // pinned staging of v_first
  bool has_columns = false;       // octant column layouts (col_cell, blk_oct, h_blk_oct) built
};

constexpr int kM2LSlots = 189;
constexpr int kM2LCols = 64;

// The 189 transfer vectors valid for each octant, in increasing v order.
const std::vector<int32_t>& octant_vlist_host();  // [8][189]
const int32_t* octant_vlist_device(hipStream_t s);

// Sorts the points by their Morton codes at `sort_level` (stable, so ties
// keep input order) and histograms, on the device into t.open_hist, the
// coarsest level at which each point opens a new cell. Nothing is read back.
void sort_tree_points(Tree& t, const double* d_pts, int64_t n, const Box& box, int sort_level,
                      TreeWorkspace& ws, hipStream_t s);

// Occupied cells per level 0..sort_level from a downloaded open_hist.
std::vector<int64_t> level_counts(const std::vector<int64_t>& open_hist);

// Builds leaf level `level` (<= t.sort_level) and every coarser level from the
// sorted points; counts[l] = occupied cells at level l. The points stay
// sorted by their codes at the sort level.
void finish_tree(Tree& t, int level, const std::vector<int64_t>& counts, TreeWorkspace& ws,
                 hipStream_t s);

// sort_tree_points + finish_tree at one level (one host round trip).
void build_tree(Tree& t, const double* d_pts, int64_t n, const Box& box, int level,
                TreeWorkspace& ws, hipStream_t s);
// P2P and M2L lists (pair-major M2L groups included); with_columns also
// builds the octant column layout that the compressed M2L path uses.
void build_lists(const Tree& trg, const Tree& src, Lists& lists, TreeWorkspace& ws, hipStream_t s,
                 bool with_columns);

}  // namespace pam
