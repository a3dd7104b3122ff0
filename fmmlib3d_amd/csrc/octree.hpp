// Host-side adaptive octree with the structure of fmmlib3d's d3tstrcr
// (src/d3tstrcr.f:24-259): same root cube, same split rule (more than nbox
// sources or more than nbox targets), same child order and stable partition,
// same breadth-first box numbering and the same interaction-list sets
// (lists 1-4).  Built level-synchronously with OpenMP and flattened to CSR so
// the device receives it in one upload.
#pragma once
#include <cstdint>
#include <vector>

namespace fmm3d {

struct Csr {
  std::vector<int> off;  // nboxes+1
  std::vector<int> idx;  // box ids
  int count(int b) const { return off[b + 1] - off[b]; }
};

struct Octree {
  int nboxes = 0, nlev = 0;
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

  bool is_leaf(int b) const { return nchild[b] == 0; }
  // empties every array but keeps its storage (repeated builds reuse it)
  void reset();
  double side(int l) const;
};

// Returns 0 on success, 128 when ns < 1 (the value d3tstrcr reports,
// d3tstrcr.f:141-145).  Coordinates are (n,3) row-major doubles.
int build_octree(const double* src, int ns, const double* trg, int nt, int nbox, Octree& t);

}  // namespace fmm3d
