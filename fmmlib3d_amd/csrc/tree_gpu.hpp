// Device octree builder (see tree_gpu.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "octree.hpp"
#include "runtime.hpp"

namespace fmm3d {

// Builds the tree of build_octree from device coordinates ((n,3) row-major
// doubles) on stream s.  The permutations (sorted position -> original
// index) are left in device buffers of ctx's arena (*sperm, *tperm; *tperm
// is null when nt = 0); t.src_perm and t.trg_perm stay empty.  With
// lists = false the interaction lists are left for tree_lists.  Returns 0
// or 128 (ns < 1).
int build_octree_device(Context& ctx, hipStream_t s, const double* src, int ns, const double* trg, int nt,
                        int nbox, Octree& t, const int** sperm, const int** tperm, bool lists = true);

// false when FMM3D_TREE=host selects the host builder for the FMM drivers
bool device_tree_enabled();

}  // namespace fmm3d
