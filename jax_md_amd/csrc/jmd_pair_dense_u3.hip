// Instantiations of the Dense-list pair dispatch (jmd_pair_dense.hpp) for
// 3-D potentials over contiguous [n, 3] positions.
#include "jmd_pair_dense.hpp"

namespace jmd {
JMD_DENSE_INSTANTIATE(3, 3, JMD_LENNARD_JONES)
JMD_DENSE_INSTANTIATE(3, 3, JMD_SOFT_SPHERE)
}  // namespace jmd
