// CPU check of the tracer's lane groups (common.hpp group_at, the layout behind group_of): for
// every workgroup size the kernels use and every front length a group round takes (1 .. block/4),
// each entry is served by exactly one group whose lanes lie in one wave, whose first lane (the one
// that dequeues the entry and broadcasts it) has g == 0, and whose spans cover neighbours 0..25
// exactly once with at most kSpan each; no thread of an entry index >= n touches the front, and
// every index a thread reads is below n. Prints "ok" or the first violation.
#include <cstdio>
#include <vector>

#include "../../brook_amd/csrc/common.hpp"

using brook::Group;
using brook::group_at;
using brook::kSpan;

int main() {
  for (int block : {256, 512, 1024}) {
    for (int n = 1; n <= block / 4; ++n) {
      std::vector<int> cover(size_t(n) * 26, 0);
      for (int tid = 0; tid < block; ++tid) {
        const Group q = group_at(n, block, tid);
        if (q.G < 4 || q.G > 32 || (q.G & (q.G - 1)) || n * q.G > block) {
          std::printf("block %d n %d tid %d: lanes per entry %d\n", block, n, tid, q.G);
          return 1;
        }
        if (q.i >= n) continue;  // idle thread: reads nothing, appends nothing
        const int lane = tid & 63, lead_tid = tid - q.g;
        if (q.lead != lane - q.g || lead_tid / 64 != tid / 64 || group_at(n, block, lead_tid).g != 0 ||
            group_at(n, block, lead_tid).i != q.i) {
          std::printf("block %d n %d tid %d: group lead lane %d outside the entry's wave\n", block, n, tid, q.lead);
          return 1;
        }
        if (q.cnt > kSpan || q.kb < 0) {
          std::printf("block %d n %d tid %d: span %d from %d\n", block, n, tid, q.cnt, q.kb);
          return 1;
        }
        for (int j = 0; j < q.cnt; ++j) {
          const int k = q.kb + j;
          if (k >= 26) {
            std::printf("block %d n %d tid %d: neighbour %d\n", block, n, tid, k);
            return 1;
          }
          cover[size_t(q.i) * 26 + k] += 1;
        }
      }
      for (size_t e = 0; e < cover.size(); ++e)
        if (cover[e] != 1) {
          std::printf("block %d n %d: entry %zu neighbour %zu covered %d times\n", block, n, e / 26, e % 26, cover[e]);
          return 1;
        }
    }
  }
  std::printf("ok\n");
  return 0;
}
