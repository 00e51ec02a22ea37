// Instantiations of the Dense-list pair dispatch (jmd_pair_dense.hpp) for
// 2-D potentials.
#include "jmd_pair_dense.hpp"

namespace jmd {
JMD_DENSE_INSTANTIATE(2, 2, JMD_LENNARD_JONES)
JMD_DENSE_INSTANTIATE(2, 2, JMD_SOFT_SPHERE)
}  // namespace jmd
