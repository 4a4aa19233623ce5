// 1-D vertex-block partition of the reference (src/partition.h:69-161):
// width = floor(N / npes), the last PE takes the remainder (max_width), and
// fewer vertices than PEs gives one vertex to each of the first N PEs.
#ifndef DMM_HOST_PARTITION_H_
#define DMM_HOST_PARTITION_H_

#include <cstdint>

namespace dmm {

struct Partition {
  int64_t N = 0;
  int pe = 0, npes = 1;
  bool small = false;
  int64_t start = 0, end = 0, partition_width = 0, max_width = 0;

  Partition() {}
  Partition(int64_t num_nodes, int pe_, int npes_) : N(num_nodes), pe(pe_), npes(npes_) {
    if (num_nodes < npes) {  // src/partition.h:83-94
      small = true;
      if (pe < num_nodes) {
        partition_width = 1;
        start = pe;
        end = pe + 1;
      } else {
        partition_width = 0;
        start = 0;
        end = 0;
      }
      max_width = 1;
    } else {  // src/partition.h:95-105
      partition_width = N / npes;
      max_width = N - (npes - 1) * partition_width;
      start = partition_width * pe;
      end = (pe == npes - 1) ? N : start + partition_width;
    }
  }

  // recv (src/partition.h:109-119): owning PE of a global vertex.
  int recv(int64_t node) const {
    if (small) return static_cast<int>(node);
    int64_t r = node / partition_width;
    return r >= npes ? npes - 1 : static_cast<int>(r);
  }

  // local_pos (src/partition.h:123-133).
  int64_t local_pos(int64_t node) const {
    if (small) return 0;
    int64_t r = node / partition_width;
    if (r >= npes) return node - (npes - 1) * partition_width;
    return node % partition_width;
  }

  int64_t global_pos(int64_t lp) const { return start + lp; }
};

}  // namespace dmm

#endif  // DMM_HOST_PARTITION_H_
