// clusters.h — the k-means light clustering shared by the host build (host_build.cpp) and the
// build whose nearest-centroid step runs on the GPU (cluster_build.hip).
#pragma once
#include <cstdint>
#include <functional>

#include "../../include/restir_abi.h"

namespace rki {

// Writes assign[l] = the index of the centroid nearest to light l (centroids c[3k] as x, y, z;
// distance (dx*dx + dy*dy) + dz*dz in fp32; ties to the lower index). Returns 0 or a
// RESTIR_E_* code.
using NearestStep = std::function<int(const float *c, uint32_t *assign)>;

// Seeded k-means (a PCG32 partial Fisher-Yates draw of k distinct lights as initial centroids,
// `iterations` Lloyd steps with centroid sums in double in light order, a final assignment),
// then the lights grouped by cluster in index order and one Vose table per cluster.
int build_light_clusters(const restir_point_light *points, uint32_t n, uint32_t k, uint64_t seed,
                         uint32_t iterations, const NearestStep &nearestStep, int32_t *lightCluster,
                         float *lightPWithin, int32_t *clusterBegin, int32_t *clusterCount,
                         int32_t *clusterLights, restir_alias_table_column *clusterAlias, float *centroidsOut);

}  // namespace rki
