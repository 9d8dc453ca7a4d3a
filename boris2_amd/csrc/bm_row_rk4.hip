// Row kernels of the RK4 stages.
#include "bm_row.h"

int bm::launch_row_rk4(const bm_demag_s* d, const StageArgs& a, int step, float scale, hipStream_t st)
{
	switch (step) {
	case 0: return launch_row_t<M_RK4, 0, true>(d, a, scale, st);
	case 1: return launch_row_t<M_RK4, 1, true>(d, a, scale, st);
	case 2: return launch_row_t<M_RK4, 2, true>(d, a, scale, st);
	case 3: return launch_row_t<M_RK4, 3, true>(d, a, scale, st);
	}
	return fail(BM_ERR_ARG, "bad RK4 stage");
}
