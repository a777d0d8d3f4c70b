// Tree metadata of a finished reconstruction (tree_metadata.py:95-153).
#pragma once
#include "psr_result.h"

#include <vector>

namespace psr {

struct TreeMetadata {
  struct Level {
    int local_depth = 0;           // -2 and -1 are the padding levels above the root
    DArray<int64_t> keys;          // global lattice coordinates (local + 2^(d+1)), tree-key encoded, sorted
    DArray<uint8_t> flags, ghost;  // per node
    DArray<int> parents;           // index in the previous level, -1 at the top
    DArray<int> children;          // 8 per node, indices in the next level (x bit first), -1 if absent
  };
  std::vector<Level> levels;
};

template <class T> void build_tree_metadata(Result<T>& r, TreeMetadata& meta, hipStream_t s);

}  // namespace psr
