// CSVDAG cross-level merge (geom_octree_extension.cpp:1192-1544).
#include "svd_internal.hpp"

namespace svd {

void cross_level_merge(svd_octree* o) {
  (void)o;
  throw Error{SVD_ERR_STATE, "cross-level merge is not available in this build"};
}

}  // namespace svd
