// Avocado protection (fix_avocados): host restatement of the reference's correction.
#pragma once
#include <cstddef>
#include <cstdint>
#include <functional>
#include <vector>

#include "common.hpp"

namespace brook {

// Corrects `labels` (compact ids below `table`, Fortran order over d) and `distance` (their DBF)
// in place; `refresh(labels, distance)` recomputes the DBF after a pass that changed labels.
// `keys`: with a voxel graph, Kimimaro's id of each compact id (ascending; the rank of its graph
// component among all of them), which its candidate and change sets hold (avocado.cpp:85-91);
// empty: the ids themselves. Returns the number of passes.
int avocado_correct(std::vector<uint32_t> &labels, std::vector<float> &distance, Dims d, size_t table, double detection,
                    const std::function<void(const std::vector<uint32_t> &, std::vector<float> &)> &refresh,
                    const std::vector<int64_t> &keys = {});

// Ids numbered by first voxel in flat order (fastremap.renumber); mapping[new id] = the input label
// (`through` of the pre-correction id in `original`) under the last run start of the new id;
// roots[new id - 1] = its first voxel. Returns the number of ids.
size_t avocado_renumber(std::vector<uint32_t> &labels, const std::vector<uint32_t> &original,
                        const std::vector<int64_t> &through, size_t table, std::vector<int64_t> &mapping,
                        std::vector<int64_t> &roots);

}  // namespace brook
