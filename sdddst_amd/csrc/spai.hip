// Learned sparse approximate inverse: placeholder until the network is
// installed (sdddst_spai_set_weights); selecting it before then is an error.
#include "context.h"
#include "solvers.h"

namespace sdamd {

int spai_factor(Context &c)
{
    SD_CHECK(c.spai != nullptr, SDDDST_ERR_STATE, "learned preconditioner selected but no weights are installed");
    return SDDDST_OK;
}

int spai_apply(Context &c, const double *, double *)
{
    SD_CHECK(c.spai != nullptr, SDDDST_ERR_STATE, "learned preconditioner selected but no weights are installed");
    return SDDDST_ERR_UNSUPPORTED;
}

}  // namespace sdamd
