// hs_uts.hip -- UTS work-stealing (in progress).
#include "hs_host.h"
extern "C" {
int hs_uts_ref_run(hs_ctx *, int, const uint32_t *, const int32_t *, const int32_t *, int,
                   hs_uts_result *) { return HS_EUNSUPPORTED; }
int hs_uts_geo_run(hs_ctx *, int, double, int, int, hs_uts_result *) { return HS_EUNSUPPORTED; }
int hs_uts_serve(hs_ctx *, const hs_uts_request *, int, const hs_serve_cfg *, hs_uts_response *,
                 hs_serve_summary *) { return HS_EUNSUPPORTED; }
}
