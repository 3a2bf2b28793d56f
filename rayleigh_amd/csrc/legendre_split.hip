// Split-precision (BF16 slice) Legendre contractions.
#include "raysht_internal.h"

namespace raysht {

int split_plan_init(raysht_plan *) { return RAYSHT_EUNSUPPORTED; }
void split_plan_free(raysht_plan *) {}
int launch_stp_split(raysht_plan *, const double *, double *, hipStream_t) { return RAYSHT_EUNSUPPORTED; }
int launch_pts_split(raysht_plan *, const double *, double *, hipStream_t) { return RAYSHT_EUNSUPPORTED; }

}  // namespace raysht
