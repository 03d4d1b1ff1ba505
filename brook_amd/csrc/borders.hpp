#pragma once
#include <array>
#include <vector>

#include "stages.hpp"

namespace brook {

using Point = std::array<int64_t, 3>;

// Per component id (count + 1 lists): the border target points in set order; the last one is
// the component's root.
std::vector<std::vector<Point>> border_targets(Context &ctx, const void *cc, int cc_bytes, Dims d, int64_t count,
                                               const float aniso[3]);

}  // namespace brook
