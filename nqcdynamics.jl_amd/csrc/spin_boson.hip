// spin_boson.hip -- system-bath kernels (placeholder until the bath kernel lands).
#include "spin_boson.h"

namespace nqcd {

int32_t sb_sample_initial(nqcd_handle_t* h, const nqcd_sampler_desc*, uint64_t) {
    return fail(h, NQCD_ERR_UNSUPPORTED, "spin-boson kernels are not built");
}
int32_t sb_run(nqcd_handle_t* h, double, int64_t, const nqcd_term_desc*, uint64_t) {
    return fail(h, NQCD_ERR_UNSUPPORTED, "spin-boson kernels are not built");
}

}  // namespace nqcd
