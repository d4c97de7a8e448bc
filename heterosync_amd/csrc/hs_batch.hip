// hs_batch.hip -- batched persistent syncPrims (in progress).
#include "hs_host.h"
extern "C" {
int hs_batch_run(hs_ctx *, const hs_batch_instance *, int, const hs_batch_cfg *,
                 int32_t *const *, int64_t *, hs_batch_result *) { return HS_EUNSUPPORTED; }
}
