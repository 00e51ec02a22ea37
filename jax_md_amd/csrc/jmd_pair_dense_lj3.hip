// Instantiations of the Dense-list pair dispatch (jmd_pair_dense.hpp) for
// 3-D Lennard-Jones over 16-byte position rows.
#include "jmd_pair_dense.hpp"

namespace jmd {
JMD_DENSE_INSTANTIATE(3, 4, JMD_LENNARD_JONES)
}  // namespace jmd
