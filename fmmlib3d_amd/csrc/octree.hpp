// Host-side adaptive octree with the structure of fmmlib3d's d3tstrcr
// (src/d3tstrcr.f:24-259): same root cube, same split rule (more than nbox
// sources or more than nbox targets), same child order and stable partition,
// same breadth-first box numbering and the same interaction-list sets
// (lists 1-4).  Built level-synchronously with OpenMP and flattened to CSR so
// the device receives it in one upload.
#pragma once
#include <cstddef>
#include <cstdint>
#include <vector>

namespace fmm3d {

struct Csr {
  std::vector<int> off;  // nboxes+1
  std::vector<int> idx;  // box ids
  int count(int b) const { return off[b + 1] - off[b]; }
};

struct Octree {
  int nboxes = 0, nlev = 0, ns = 0, nt = 0;
  double center[3] = {0, 0, 0}, size = 0;
  // per box
  std::vector<int> level, parent, nchild, child;  // child: 8 per box, -1 unused
  std::vector<int> ijk;                           // 3 per box, 1-based like the reference
  std::vector<int> isrc, nsrc, itrg, ntrg;        // 0-based starts in sorted order
  std::vector<double> bcenter;                    // 3 per box
  std::vector<int> ladd;                          // first box of level l at ladd[l], nlev+2 entries
  // permutations: sorted position -> original index
  std::vector<int> src_perm, trg_perm;
  // interaction lists (index 0..3 = reference lists 1..4), ascending per box
  Csr list[4];
  // Forest of several trees (merge_forest): per box the index of its
  // (tree, level) scale and 1/side per index; empty for a single tree, whose
  // scale is 1/side(level)
  std::vector<int> scale_id;
  std::vector<double> scale_inv;

  bool is_leaf(int b) const { return nchild[b] == 0; }
  // every leaf that holds points lies on one level (then lists 3 and 4 are
  // empty: no interactions between boxes of different sizes)
  bool leaves_on_one_level() const {
    int lv = -1;
    for (int b = 0; b < nboxes; b++) {
      if (nchild[b] != 0 || nsrc[b] + ntrg[b] == 0) continue;
      if (lv < 0) lv = level[b];
      else if (level[b] != lv) return false;
    }
    return true;
  }
  // empties every array but keeps its storage (repeated builds reuse it)
  void reset();
  double side(int l) const;
};

// Returns 0 on success, 128 when ns < 1 (the value d3tstrcr reports,
// d3tstrcr.f:141-145).  Coordinates are (n,3) row-major doubles in host memory.
// With lists = false the interaction lists are left for tree_lists.
int build_octree(const double* src, int ns, const double* trg, int nt, int nbox, Octree& t, bool lists = true);

// ---- pieces of the level-synchronous build shared with the device builder
constexpr int kTreeMaxLevel = 100;

// One pass over the points resolves up to D levels below the frontier boxes.
struct TreePass {
  std::vector<int> frontier;  // boxes split in this pass
  int L = 0;                  // their level
  int D = 1, Dfull = 1, Dk = 1;  // digits per code, digits resolved
  int NB = 8, CT = 4, NS = 16;
  std::vector<double> ctab;   // centres: [f][dim][depth][index], width CT
  std::vector<int> hs, ht;    // per frontier box: NB+1 exclusive prefix counts of the codes
  std::vector<char> splits;   // per frontier box: NS flags of the sub-boxes that split further
  int posd[8] = {0, 0, 0, 0, 0, 0, 0, 0};
};

// root box from the bounding box of all points; frontier = {0} if it splits
void tree_root(Octree& t, const double lo[3], const double hi[3], int ns, int nt, int nbox,
               std::vector<int>& frontier);
// digits, centre table and empty histograms for ps.frontier at level L
void tree_pass_setup(const Octree& t, int L, TreePass& ps);
// with ps.hs/ps.ht filled: create the sub-boxes, mark the splits, return the
// boxes to split in the next pass in `next`; sets ps.Dk
void tree_pass_expand(Octree& t, int nbox, TreePass& ps, std::vector<int>& next);
// code of a point truncated at the depth where its box stops splitting
inline uint32_t tree_eff_code(const TreePass& ps, int f, uint32_t code) {
  const int Dk = ps.Dk;
  const uint32_t c = code >> (3 * (ps.Dfull - Dk));
  const char* sp = &ps.splits[(size_t)f * ps.NS];
  for (int d = 1; d < Dk; d++) {
    const uint32_t pre = c >> (3 * (Dk - d));
    if (!sp[ps.posd[d] + pre]) return pre << (3 * (Dk - d));
  }
  return c;
}
// Disjoint union of finished trees (lists built) as one Octree, for running
// several small problems through one set of launches: boxes ordered by level,
// then tree, then their own order (so every tree keeps its numbering order
// and a level's boxes stay contiguous in ladd); point ranges and
// permutations shifted by src_off / trg_off (the trees' points concatenated
// in tree order); lists renumbered; scale_id / scale_inv carry each tree's
// box sizes.
void merge_forest(const std::vector<const Octree*>& trees, Octree& f);

// colleagues and lists 1-4 from the finished box structure
void tree_lists(Octree& t);
// the same in two parts on one thread: colleagues, lists 1 and 3; then lists 2
// and 4 (needs the first part)
void tree_lists_near(Octree& t);
void tree_lists_far(Octree& t);

}  // namespace fmm3d
