// Row kernels of the RKDP54 stages (stage 0 has no next-evaluation forward transform: the
// advance that follows it runs in the forward x pass of the convolution).
#include "bm_row.h"

int bm::launch_row_rkdp(const bm_demag_s* d, const StageArgs& a, int step, float scale, hipStream_t st)
{
	switch (step) {
	case 0: return launch_row_t<M_RKDP, 0, false>(d, a, scale, st);
	case 1: return launch_row_t<M_RKDP, 1, true>(d, a, scale, st);
	case 2: return launch_row_t<M_RKDP, 2, true>(d, a, scale, st);
	case 3: return launch_row_t<M_RKDP, 3, true>(d, a, scale, st);
	case 4: return launch_row_t<M_RKDP, 4, true>(d, a, scale, st);
	case 5: return launch_row_t<M_RKDP, 5, true>(d, a, scale, st);
	}
	return fail(BM_ERR_ARG, "bad RKDP54 stage");
}
