// placeholder, replaced by the mixed-precision engine
#include "cx_internal.hpp"
namespace cx {
void knn_mixed(TreeState &t, const int64_t *d_queries, int64_t nq, int32_t k, int32_t predecessor,
               int64_t *d_idx, float *d_dist, float *d_err)
{
    throw Error(CX_ERR_UNSUPPORTED, "mixed engine not built");
}
} // namespace cx
