// method = MR_OP_NUFFT engine (type-2 NUFFT, per-readout Gram multiply,
// type-1 NUFFT per coil); see mr_op.hip.

#include "mr_grid.hpp"

struct mr_nufft_eng {
    int unused;
};

int mr_nufft_eng_create(const mr_config *, const mr_complex *, const float *, int64_t, int64_t,
                        double, mr_nufft_eng **out)
{
    *out = nullptr;
    MR_FAIL(MR_ERR_UNSUPPORTED, "mr_op_create: the NUFFT method is not available in this build");
}

int mr_nufft_eng_execute(mr_nufft_eng *, const mr_complex *, const mr_complex *, mr_complex *)
{
    MR_FAIL(MR_ERR_UNSUPPORTED, "the NUFFT method is not available in this build");
}

void mr_nufft_eng_destroy(mr_nufft_eng *e) { delete e; }
