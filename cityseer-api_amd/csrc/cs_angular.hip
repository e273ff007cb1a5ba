// cs_angular.hip -- simplest-path (angular) centrality on the dual graph.
#include "cs_internal.hpp"

namespace cs {

void launch_simplest(Graph &g, DevGraph &dg, DevEdges &de, const CallSpec &cs,
                     const uint32_t *d_sources, const float *d_wt, uint32_t n_sources,
                     const uint8_t *d_eligible, double *d_acc, uint32_t *d_reach,
                     hipStream_t stream, cs_stats &st) {
  (void)g; (void)dg; (void)de; (void)cs; (void)d_sources; (void)d_wt; (void)n_sources;
  (void)d_eligible; (void)d_acc; (void)d_reach; (void)stream; (void)st;
  fail(CS_EINVAL, "centrality_simplest is not available on the GPU path yet");
}

}  // namespace cs
