// Row kernels of the TEuler (Heun) and AHeun stages.
#include "bm_row.h"

int bm::launch_row_heun(const bm_demag_s* d, const StageArgs& a, int mcode, int step, float scale, hipStream_t st)
{
	if (mcode == M_TEULER) return step ? launch_row_t<M_TEULER, 1, true>(d, a, scale, st) : launch_row_t<M_TEULER, 0, true>(d, a, scale, st);
	if (mcode == M_AHEUN) return step ? launch_row_t<M_AHEUN, 1, true>(d, a, scale, st) : launch_row_t<M_AHEUN, 0, true>(d, a, scale, st);
	return fail(BM_ERR_ARG, "bad Heun method");
}
