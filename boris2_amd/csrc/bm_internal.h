// Internal declarations shared by the library's translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <string>

#include "boris_mi355x.h"

namespace bm {

void set_error(const std::string& msg);
int fail(int code, const std::string& msg);
int hip_fail(hipError_t e, const char* what);

#define BM_HIP(call) do { hipError_t e_ = (call); if (e_ != hipSuccess) return ::bm::hip_fail(e_, #call); } while (0)

int device_cu_count();

// Launch timing: when enabled (bm_prof_begin), every kernel launch of the hot path is
// bracketed by HIP events on its own stream and tagged with its kind.
enum ProfKind { PK_FWD_X = 0, PK_Y_FWD, PK_Z, PK_Y_INV, PK_INV_X, PK_Y2D, PK_STAGE, PK_ADVANCE, PK_ROW, PK_FWD_ADV, PK_COUNT };
void prof_mark_begin(int kind, hipStream_t st);
void prof_mark_end(int kind, hipStream_t st);

} // namespace bm

struct bm_demag_s {
	int nx, ny, nz;
	int Nx, Ny, Nz, NC, NYH, NZH;
	int Sp;                   // line pitch of the spectra (complex elements, >= Ny)
	double h[3];
	bool is2d;
	float2* S = nullptr;      // spectra: 3 * nz * NC * Ny complex
	float* K = nullptr;       // kernels (see bm_demag_copy_kernels)
	long long kcount = 0;
	double* ebuf = nullptr;   // energy partials (one per P5 work item)
	int ecap = 0;
	long long bytes = 0;
	int ncu = 256;
};

namespace bm {
// H_demag(M) into planar output H3n (3 planes of n floats)
int demag_planar(bm_demag_s* d, const float* M, float* H3n, hipStream_t st);
// the convolution's passes, for the integrator's fused evaluation
int demag_fwd_x(bm_demag_s* d, const float* M, hipStream_t st);
int demag_fwd_x_advance(bm_demag_s* d, const float* M, const float* E0, float* sM1, float* Mout, float dT5, hipStream_t st);
int demag_mid(bm_demag_s* d, hipStream_t st);
}
