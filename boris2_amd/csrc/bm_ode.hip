// LLG integration of one ferromagnetic mesh on MI355X: stage kernels and the host-side
// step-control state machine.
//
// With demag enabled, a field evaluation is the convolution's middle passes (y / z FFTs and
// the kernel product) followed by one fused row kernel — x-direction inverse FFT of the
// demag field, the per-cell field + LLG + stage update (bm_stage.h), and the x-direction
// forward FFT of the stage's output magnetisation, which is the next evaluation's input.
// Without demag a plain per-cell stage kernel runs.
// The host drives stages exactly as ODECommon_Base::IterateCUDA does
// (Boris/DiffEq_CommonBase_IterateCUDA.cpp:13-936): one lte read-back per adaptive step,
// SetAdaptiveTimeStep in FP64 on the host (Boris/DiffEq_CommonBase_Iterate.cpp:15-59),
// restore + re-prime on rejection.
#include "bm_internal.h"
#include "bm_stage.h"
#include "bm_row.h"

#include <cmath>
#include <cstring>
#include <algorithm>
#include <cstdlib>
#include <vector>

namespace {

using namespace bm;

// One cell of a stage sweep: its demag field (planar convolution output, or the evaluation
// speedup's extrapolation / save) and the fused field + LLG + stage update.
template <int METHOD, int STEP>
__device__ __forceinline__ void stage_idx(const StageArgs& a, F3 Ha, size_t idx, size_t n, float& lte_loc)
{
	const int i = (int)(idx % a.nx), j = (int)((idx / a.nx) % a.ny), k = (int)(idx / ((size_t)a.nx * a.ny));
	F3 Hd{ 0.f, 0.f, 0.f };
	if (a.nsp || a.Hsave) {
		// self coefficient of the cell: the mesh's, or its layer's in a layer stack
		F3 sc{ a.selfc[0], a.selfc[1], a.selfc[2] };
		if (a.layered) sc = F3{ a.lay->selfc[k][0], a.lay->selfc[k][1], a.lay->selfc[k][2] };
		const F3 m = ld3(a.Min, idx);
		if (a.nsp) {
			// Demag_EvalSpeedup_AddExtrapField_AddSelf: saved fields in slot order, then self term
			Hd = a.ca[0] * F3{ a.Hs[0][idx], a.Hs[0][n + idx], a.Hs[0][2 * n + idx] };
#pragma unroll
			for (int q = 1; q < 6; q++)
				if (q < a.nsp) Hd = Hd + a.ca[q] * F3{ a.Hs[q][idx], a.Hs[q][n + idx], a.Hs[q][2 * n + idx] };
			Hd = Hd + F3{ sc.x * m.x, sc.y * m.y, sc.z * m.z };
		}
		else {
			// Demag_EvalSpeedup_AddField_SubSelf: keep the convolution minus the self term
			Hd = F3{ a.Hd[idx], a.Hd[n + idx], a.Hd[2 * n + idx] };
			a.Hsave[idx] = Hd.x - sc.x * m.x;
			a.Hsave[n + idx] = Hd.y - sc.y * m.y;
			a.Hsave[2 * n + idx] = Hd.z - sc.z * m.z;
		}
	}
	else if (a.modules & BM_MOD_DEMAG) Hd = F3{ a.Hd[idx], a.Hd[n + idx], a.Hd[2 * n + idx] };
	stage_cell<METHOD, STEP>(a, Ha, idx, i, j, k, Hd, lte_loc);
}

// One stage over a mesh (all the block's cells).
template <int METHOD, int STEP>
__device__ __forceinline__ void stage_sweep(const StageArgs& a)
{
	const F3 Ha = stage_Ha(a);
	const size_t n = (size_t)a.nx * a.ny * a.nz;
	const long long nchunk = (long long)((n + kBS - 1) / kBS);
	float lte_loc = 0.f;
	// a.map 0: grid-stride over chunks; 1: one chunk per block, each XCD a contiguous eighth;
	// 2: grid-stride sweeps of gridDim chunks, each XCD a contiguous eighth of every sweep
	for (long long r = 0;; r++) {
		long long chunk;
		if (a.map == 1) { if (r) break; chunk = xcd_chunk(blockIdx.x, gridDim.x); }
		else if (a.map == 2) chunk = r * gridDim.x + xcd_chunk(blockIdx.x, gridDim.x);
		else chunk = r * gridDim.x + blockIdx.x;
		if (chunk >= nchunk) { if (a.map == 1) break; if (r * gridDim.x >= nchunk) break; continue; }
		const size_t idx = (size_t)chunk * kBS + threadIdx.x;
		if (idx >= n) continue;
		stage_idx<METHOD, STEP>(a, Ha, idx, n, lte_loc);
	}
	if (stage_has_lte<METHOD, STEP>()) block_max_to(lte_loc, a.lte);
}

// Batched: one block per chunk of kBS cells over all meshes, mesh b owning the flat chunks
// choff[b] .. choff[b + 1] - 1 (so no block is sized by another mesh); blocks take
// XCD-contiguous chunks (gridDim.x a multiple of 8), so neighbouring chunks of a mesh, whose
// rows the exchange stencil shares, land on one L2.  The mesh's arguments are read in place
// from its table entry (a local copy of the large argument set would go to scratch).
template <int METHOD, int STEP>
__device__ __forceinline__ void stage_chunk(const StageArgs* __restrict__ bt, const int* __restrict__ choff, int nb)
{
	const int c = xcd_chunk(blockIdx.x, gridDim.x);
	if (c >= choff[nb]) return;                       // uniform over the block
	int lo = 0, hi = nb - 1;                          // last mesh with choff[mesh] <= c
	while (lo < hi) { const int mid = (lo + hi + 1) >> 1; if (choff[mid] <= c) lo = mid; else hi = mid - 1; }
	const StageArgs& a = bt[lo];
	if (a.ctl && (STEP == 0 ? !a.ctl->active : a.ctl->skip)) return;   // uniform over the block
	const size_t n = (size_t)a.nx * a.ny * a.nz;
	const size_t idx = (size_t)(c - choff[lo]) * kBS + threadIdx.x;
	float lte_loc = 0.f;
	if (idx < n) stage_idx<METHOD, STEP>(a, stage_Ha(a), idx, n, lte_loc);
	if (stage_has_lte<METHOD, STEP>()) block_max_to(lte_loc, a.lte);
}

template <int METHOD, int STEP>
__global__ void __launch_bounds__(kBS) k_stage(StageArgs a, const StageArgs* __restrict__ bt, const int* __restrict__ choff, int nb)
{
	if (bt) stage_chunk<METHOD, STEP>(bt, choff, nb);    // batched: flat chunks of all meshes
	else stage_sweep<METHOD, STEP>(a);
}

// RKDP54 advance after the step size is fixed: sM1 = M; M1 = M + E0 dT/5 (into the
// second magnetisation buffer, which stage 1 then reads).
__global__ void __launch_bounds__(kBS) k_rkdp_advance(const float* __restrict__ M, float* __restrict__ M1, float* __restrict__ sM1,
	const float* __restrict__ E0, size_t n3, float dT5)
{
	for (size_t q = blockIdx.x * (size_t)kBS + threadIdx.x; q < n3; q += (size_t)gridDim.x * kBS) {
		float m = M[q];
		sM1[q] = m;
		M1[q] = m + E0[q] * dT5;
	}
}

__global__ void k_avg_partial(const float* __restrict__ M, const float* __restrict__ Ms_map, float Ms, size_t n, double* part)
{
	__shared__ double red[3][kBS];
	double s[3] = { 0, 0, 0 };
	for (size_t i = blockIdx.x * (size_t)kBS + threadIdx.x; i < n; i += (size_t)gridDim.x * kBS) {
		const float ms = Ms_map ? Ms_map[i] : Ms;
		s[0] += (double)(M[3 * i] / ms); s[1] += (double)(M[3 * i + 1] / ms); s[2] += (double)(M[3 * i + 2] / ms);
	}
	for (int c = 0; c < 3; c++) red[c][threadIdx.x] = s[c];
	__syncthreads();
	for (int o = kBS / 2; o > 0; o >>= 1) {
		if (threadIdx.x < o) for (int c = 0; c < 3; c++) red[c][threadIdx.x] += red[c][threadIdx.x + o];
		__syncthreads();
	}
	if (threadIdx.x == 0) for (int c = 0; c < 3; c++) part[3 * blockIdx.x + c] = red[c][0];
}

} // namespace

struct bm_ode_s {
	int method, mcode;
	int nx, ny, nz;
	size_t n;
	double h[3];
	bm_fm_params p;
	bm_demag_t demag = nullptr;
	// layer stack (bm_ode_set_layers): multilayered demag and per-layer material values
	bm_mlayer_t ml = nullptr;
	bool lay_set[3] = { false, false, false };   // A, alpha, K1 given per layer
	float lay[3][bm::kMaxStackLayers] = {};
	// stacks of 3D meshes (bm_ode_set_layer_meshes): mesh index of each layer, exchange
	// spacing across planes of a mesh, per-layer constant applied fields
	bool lay_grouped = false, lay_field = false;
	int lay_grp[bm::kMaxStackLayers] = {};
	float lay_ihz2[bm::kMaxStackLayers] = {};
	float lay_Ha[bm::kMaxStackLayers][3] = {};
	bm::LayerTab* d_lay = nullptr;   // device copy of the per-layer values (upload_layers)
	float* Ms_layers = nullptr;    // per-cell Ms built from per-layer values (owned)
	float* M = nullptr;            // the mesh's magnetisation (caller-owned); holds M between steps
	float* Mb = nullptr;           // second magnetisation buffer for intermediate stages
	float* Mcur = nullptr;         // the array holding the magnetisation of the current stage
	const float* Ms_map = nullptr;
	float* Hd = nullptr;
	float* sM1 = nullptr;
	float* E[6] = { nullptr };
	float* Heff_keep = nullptr;   // internal AoS copy of the last effective field
	bool use_row = true;          // demag evaluations through the fused row kernel
	bool S_valid = false;         // the demag spectra hold the x-transform of Mcur
	bool keep_heff = false;
	unsigned* d_lte = nullptr;
	unsigned* h_lte = nullptr;    // pinned
	double* d_part = nullptr;
	double* h_part = nullptr;     // pinned
	int npart = 0;
	int ncu = 256;
	// ODE state (ODECommon_Base)
	double time = 0, stagetime = 0, dT = 0, dT_last = 0, lte = 0;
	double err_high_fail = 1e-4, dT_increase = 2, dT_min = 1e-15, dT_max = 3e-12;
	int order = 5, evalStep = 0;
	long long iteration = 0, stageiteration = 0, rejections = 0, evaluations = 0;
	bool primed = false, available = true;
	// demag evaluation speedup (DemagCUDA::UpdateField, Boris/DemagCUDA.cpp:106-468)
	int speedup = 0, num_saved = 0;       // EVALSPEEDUP_ mode = number of saved fields
	double tdm[6] = {};                   // save times
	float* Hs[6] = {};                    // saved fields minus self term, planar
	float selfc[3] = {};
	float lay_selfc[bm::kMaxStackLayers][3] = {};
	bool link_dTspeedup = true;
	double dTspeedup = 0, time_speedup = 0;
	long long convolutions = 0;
	// periodic boundary conditions (bm_ode_set_pbc): images per axis; exchange wraps around
	// along every axis with images, and the demag handle must carry the same images
	int pbc[3] = { 0, 0, 0 };
	// per-launch speedup use, read by make_args: extrapolate (sp_n saved fields) or save
	int sp_n = 0;
	float sp_ca[6] = {};
	float* sp_save = nullptr;
	// fields of other modules (bm_ode_set_extra_field): called before every stage with the
	// magnetisation that stage evaluates; its output is added to the effective field
	bm_field_fn ext_fn = nullptr;
	void* ext_user = nullptr;
	float* Hext = nullptr;
};

namespace {

int stage_grid(bm_ode_s* o) { return (int)std::max<size_t>(1, std::min<size_t>((o->n + kBS - 1) / kBS, (size_t)o->ncu * 16)); }
// stage kernels: one kBS-cell chunk per block, rounded up to a multiple of 8 blocks (XCD map)
// Block-to-cell mapping of the stage kernels (BM_STAGE_MAP overrides; 2 measured fastest).
int stage_map()
{
	static int m = -1;
	if (m < 0) { const char* e = getenv("BM_STAGE_MAP"); m = e ? atoi(e) : 2; }
	return m;
}
int stage_chunks(bm_ode_s* o)
{
	size_t nb = (o->n + kBS - 1) / kBS;
	if (stage_map() == 1) return (int)(((nb + 7) / 8) * 8);
	return (int)std::max<size_t>(8, std::min<size_t>(((nb + 7) / 8) * 8, (size_t)o->ncu * 16));
}

void applied_field(const bm_fm_params& p, double t, double H[3])
{
	for (int c = 0; c < 3; c++) H[c] = p.Ha[c];
	if (p.wave == 1) {
		double x = p.w * (t - p.t0);
		double s = (std::fabs(x) < 1e-300) ? 1.0 : std::sin(x) / x;
		for (int c = 0; c < 3; c++) H[c] += p.amp * p.dir[c] * s;
	}
	else if (p.wave == 2) {
		for (int c = 0; c < 3; c++) H[c] += p.slope[c] * t;
	}
}

StageArgs make_args(bm_ode_s* o, float dT)
{
	StageArgs a;
	std::memset(&a, 0, sizeof(a));
	a.Min = o->Mcur; a.Mout = (o->Mcur == o->M) ? o->Mb : o->M; a.Hd = o->Hd; a.sM1 = o->sM1;
	for (int q = 0; q < 6; q++) a.E[q] = o->E[q];
	a.Ms_map = o->Ms_map; a.Heff_out = o->keep_heff ? o->Heff_keep : nullptr; a.lte = o->d_lte;
	a.Hext = o->ext_fn ? o->Hext : nullptr;
	a.nx = o->nx; a.ny = o->ny; a.nz = o->nz;
	a.per = (o->pbc[0] ? 1 : 0) | (o->pbc[1] ? 2 : 0) | (o->pbc[2] ? 4 : 0);
	a.ihx2 = (float)(1.0 / (o->h[0] * o->h[0])); a.ihy2 = (float)(1.0 / (o->h[1] * o->h[1])); a.ihz2 = (float)(1.0 / (o->h[2] * o->h[2]));
	a.Ms = (float)o->p.Ms; a.A = (float)o->p.A; a.alpha = (float)o->p.alpha; a.K1 = (float)o->p.K1; a.K2 = (float)o->p.K2;
	a.grel = (float)o->p.grel;
	for (int c = 0; c < 3; c++) a.ea[c] = (float)o->p.ea[c];
	double H[3]; applied_field(o->p, o->stagetime, H);
	for (int c = 0; c < 3; c++) a.Ha[c] = (float)H[c];
	a.dT = dT;
	a.modules = o->p.modules;
	a.map = stage_map();
	if (o->sp_n || o->sp_save) {
		a.nsp = o->sp_n;
		for (int q = 0; q < 6; q++) { a.Hs[q] = o->Hs[q]; a.ca[q] = o->sp_ca[q]; }
		for (int c = 0; c < 3; c++) a.selfc[c] = o->selfc[c];
		a.Hsave = o->sp_save;
	}
	if (o->ml) {
		a.layered = 1;
		a.ihz2 = 0.f;
		a.lay_grouped = o->lay_grouped ? 1 : 0;
		a.lay_field = o->lay_field ? 1 : 0;
		a.lay = o->d_lay;
	}
	return a;
}

int launch_stage(bm_ode_s* o, int step, float dT, hipStream_t st)
{
	StageArgs a = make_args(o, dT);
	dim3 g(stage_chunks(o)), b(kBS);
	bm::prof_mark_begin(bm::PK_STAGE, st);
	switch (o->mcode * 8 + step) {
	case M_RKDP * 8 + 0: hipLaunchKernelGGL((k_stage<M_RKDP, 0>), g, b, 0, st, a, (const StageArgs*)nullptr, (const int*)nullptr, 0); break;
	case M_RKDP * 8 + 1: hipLaunchKernelGGL((k_stage<M_RKDP, 1>), g, b, 0, st, a, (const StageArgs*)nullptr, (const int*)nullptr, 0); break;
	case M_RKDP * 8 + 2: hipLaunchKernelGGL((k_stage<M_RKDP, 2>), g, b, 0, st, a, (const StageArgs*)nullptr, (const int*)nullptr, 0); break;
	case M_RKDP * 8 + 3: hipLaunchKernelGGL((k_stage<M_RKDP, 3>), g, b, 0, st, a, (const StageArgs*)nullptr, (const int*)nullptr, 0); break;
	case M_RKDP * 8 + 4: hipLaunchKernelGGL((k_stage<M_RKDP, 4>), g, b, 0, st, a, (const StageArgs*)nullptr, (const int*)nullptr, 0); break;
	case M_RKDP * 8 + 5: hipLaunchKernelGGL((k_stage<M_RKDP, 5>), g, b, 0, st, a, (const StageArgs*)nullptr, (const int*)nullptr, 0); break;
	case M_RK4 * 8 + 0: hipLaunchKernelGGL((k_stage<M_RK4, 0>), g, b, 0, st, a, (const StageArgs*)nullptr, (const int*)nullptr, 0); break;
	case M_RK4 * 8 + 1: hipLaunchKernelGGL((k_stage<M_RK4, 1>), g, b, 0, st, a, (const StageArgs*)nullptr, (const int*)nullptr, 0); break;
	case M_RK4 * 8 + 2: hipLaunchKernelGGL((k_stage<M_RK4, 2>), g, b, 0, st, a, (const StageArgs*)nullptr, (const int*)nullptr, 0); break;
	case M_RK4 * 8 + 3: hipLaunchKernelGGL((k_stage<M_RK4, 3>), g, b, 0, st, a, (const StageArgs*)nullptr, (const int*)nullptr, 0); break;
	case M_TEULER * 8 + 0: hipLaunchKernelGGL((k_stage<M_TEULER, 0>), g, b, 0, st, a, (const StageArgs*)nullptr, (const int*)nullptr, 0); break;
	case M_TEULER * 8 + 1: hipLaunchKernelGGL((k_stage<M_TEULER, 1>), g, b, 0, st, a, (const StageArgs*)nullptr, (const int*)nullptr, 0); break;
	case M_AHEUN * 8 + 0: hipLaunchKernelGGL((k_stage<M_AHEUN, 0>), g, b, 0, st, a, (const StageArgs*)nullptr, (const int*)nullptr, 0); break;
	case M_AHEUN * 8 + 1: hipLaunchKernelGGL((k_stage<M_AHEUN, 1>), g, b, 0, st, a, (const StageArgs*)nullptr, (const int*)nullptr, 0); break;
	default: return bm::fail(BM_ERR_ARG, "bad stage");
	}
	bm::prof_mark_end(bm::PK_STAGE, st);
	// every stage except RKDP54 stage 0 writes the next magnetisation into the other array
	if (!(o->mcode == M_RKDP && step == 0)) o->Mcur = (o->Mcur == o->M) ? o->Mb : o->M;
	return BM_OK;
}

} // namespace

// Batched stage launch: nb meshes, per-mesh StageArgs in d_tab, mesh b owning the flat chunks
// choff[b] .. choff[b + 1] - 1 of kBS cells (choff[nb] = nchunks); one block per chunk.
int bm::batch_stage(int mcode, int step, const StageArgs* d_tab, int nb, const int* d_choff, int nchunks, hipStream_t st)
{
	const dim3 g((unsigned)(((std::max(nchunks, 1) + 7) / 8) * 8)), b(kBS);
	StageArgs a{};
	switch (mcode * 8 + step) {
	case M_RKDP * 8 + 0: hipLaunchKernelGGL((k_stage<M_RKDP, 0>), g, b, 0, st, a, d_tab, d_choff, nb); break;
	case M_RKDP * 8 + 1: hipLaunchKernelGGL((k_stage<M_RKDP, 1>), g, b, 0, st, a, d_tab, d_choff, nb); break;
	case M_RKDP * 8 + 2: hipLaunchKernelGGL((k_stage<M_RKDP, 2>), g, b, 0, st, a, d_tab, d_choff, nb); break;
	case M_RKDP * 8 + 3: hipLaunchKernelGGL((k_stage<M_RKDP, 3>), g, b, 0, st, a, d_tab, d_choff, nb); break;
	case M_RKDP * 8 + 4: hipLaunchKernelGGL((k_stage<M_RKDP, 4>), g, b, 0, st, a, d_tab, d_choff, nb); break;
	case M_RKDP * 8 + 5: hipLaunchKernelGGL((k_stage<M_RKDP, 5>), g, b, 0, st, a, d_tab, d_choff, nb); break;
	case M_RK4 * 8 + 0: hipLaunchKernelGGL((k_stage<M_RK4, 0>), g, b, 0, st, a, d_tab, d_choff, nb); break;
	case M_RK4 * 8 + 1: hipLaunchKernelGGL((k_stage<M_RK4, 1>), g, b, 0, st, a, d_tab, d_choff, nb); break;
	case M_RK4 * 8 + 2: hipLaunchKernelGGL((k_stage<M_RK4, 2>), g, b, 0, st, a, d_tab, d_choff, nb); break;
	case M_RK4 * 8 + 3: hipLaunchKernelGGL((k_stage<M_RK4, 3>), g, b, 0, st, a, d_tab, d_choff, nb); break;
	case M_TEULER * 8 + 0: hipLaunchKernelGGL((k_stage<M_TEULER, 0>), g, b, 0, st, a, d_tab, d_choff, nb); break;
	case M_TEULER * 8 + 1: hipLaunchKernelGGL((k_stage<M_TEULER, 1>), g, b, 0, st, a, d_tab, d_choff, nb); break;
	default: return bm::fail(BM_ERR_ARG, "batched stage: RKDP54, RK4 or TEuler");
	}
	return BM_OK;
}

namespace {

// Fused evaluation: [x-R2C of Mcur unless the spectra already hold it] + y / z passes with
// the kernel product + the row kernel (x-C2R of H_demag, field + stage, x-R2C of the stage's
// output magnetisation, which the next evaluation starts from).
int eval_row(bm_ode_s* o, int step, float dT, hipStream_t st)
{
	bm_demag_s* d = o->demag;
	int rc;
	if (!o->S_valid && (rc = bm::demag_fwd_x(d, o->Mcur, st))) return rc;
	if ((rc = bm::demag_mid(d, st))) return rc;
	StageArgs a = make_args(o, dT);
	const float scale = (float)(1.0 / ((double)d->Nx * d->Ny * d->Nz));
	bm::prof_mark_begin(bm::PK_ROW, st);
	if (o->mcode == M_RKDP) rc = bm::launch_row_rkdp(d, a, step, scale, st);
	else if (o->mcode == M_RK4) rc = bm::launch_row_rk4(d, a, step, scale, st);
	else rc = bm::launch_row_heun(d, a, o->mcode, step, scale, st);
	bm::prof_mark_end(bm::PK_ROW, st);
	if (rc) return rc;
	const bool writes_m = !(o->mcode == M_RKDP && step == 0);
	o->S_valid = writes_m;
	if (writes_m) o->Mcur = (o->Mcur == o->M) ? o->Mb : o->M;
	return BM_OK;
}

// ODECommon_Base::Get_EvalStep_Time (Boris/DiffEq_CommonBase_Iterate.cpp:1536-1635): step
// start time plus dT times the tableau's c value of the stage about to run.
double eval_step_time(const bm_ode_s* o)
{
	static const double c_teuler[2] = { 0.0, 1.0 };
	static const double c_rk4[4] = { 0.0, 0.5, 0.5, 1.0 };
	static const double c_rkdp54[6] = { 0.0, 0.2, 0.3, 0.8, 8.0 / 9, 1.0 };
	switch (o->mcode) {
	case M_TEULER: case M_AHEUN: return o->time + o->dT * c_teuler[o->evalStep];
	case M_RK4: return o->time + o->dT * c_rk4[o->evalStep];
	case M_RKDP: return o->time + o->dT * c_rkdp54[o->evalStep];
	}
	return o->time;
}

// ODECommon_Base::Check_Step_Update (Boris/DiffEq_CommonBase_Iterate.cpp:1639-1657)
bool check_step_update(bm_ode_s* o)
{
	if (!o->speedup) return true;
	if (!o->link_dTspeedup) {
		const double te = eval_step_time(o);
		if (te < o->time_speedup + o->dTspeedup) return false;
		o->time_speedup = te;
		return true;
	}
	return o->available;
}

// Slot a new save replaces once all are filled: the oldest, from the order of the save
// times (Boris/DemagCUDA.cpp:176-324).
int next_slot(const bm_ode_s* o)
{
	const int p = o->speedup;
	if (p == 1) return 0;
	bool rising = true;
	for (int q = 1; q < p; q++) rising = rising && o->tdm[q] > o->tdm[q - 1];
	if (rising) return 0;
	for (int q = 0; q + 2 < p; q++) if (o->tdm[q] > o->tdm[q + 1]) return q + 1;
	return p - 1;
}

// Demag evaluation with the speedup (DemagCUDA::UpdateField, Boris/DemagCUDA.cpp:106-468):
// a full convolution (saved, minus the self term, while the ODE recommends an update or
// slots are free), or the Lagrange extrapolation of the saved fields plus the current self
// term (coefficients in FP64 on the host, rounded to FP32 as the reference's cuBReal).
int eval_speedup(bm_ode_s* o, int step, float dT, hipStream_t st)
{
	const int p = o->speedup;
	int rc;
	auto convolve = [&]() { return o->ml ? bm::mlayer_planar(o->ml, o->Mcur, o->Hd, st) : bm::demag_planar(o->demag, o->Mcur, o->Hd, st); };
	if (o->num_saved < p && !check_step_update(o)) {
		if ((rc = convolve())) return rc;
		o->convolutions++;
		return launch_stage(o, step, dT, st);
	}
	if (check_step_update(o) || o->num_saved < p) {
		int slot;
		if (o->num_saved < p) { slot = o->num_saved; o->tdm[slot] = eval_step_time(o); o->num_saved++; }
		else { slot = next_slot(o); if (p > 1) o->tdm[slot] = eval_step_time(o); }
		if ((rc = convolve())) return rc;
		o->convolutions++;
		o->sp_save = o->Hs[slot];
		rc = launch_stage(o, step, dT, st);
		o->sp_save = nullptr;
		return rc;
	}
	const double t = eval_step_time(o);
	for (int q = 0; q < 6; q++) o->sp_ca[q] = 0.f;
	if (p == 1) o->sp_ca[0] = 1.f;
	else for (int q = 0; q < p; q++) {
		double num = 1, den = 1;
		for (int r = 0; r < p; r++) if (r != q) { num *= (t - o->tdm[r]); den *= (o->tdm[q] - o->tdm[r]); }
		o->sp_ca[q] = (float)(num / den);
	}
	o->sp_n = p;
	rc = launch_stage(o, step, dT, st);
	o->sp_n = 0;
	return rc;
}

// Per-layer values of a stack into its device table (A, alpha, K1 given per layer or the
// mesh's, self coefficients, mesh grouping, fields); called whenever any of them changes.
int upload_layers(bm_ode_s* o)
{
	if (!o->ml) return BM_OK;
	if (!o->d_lay && hipMalloc(&o->d_lay, sizeof(bm::LayerTab)) != hipSuccess) { o->d_lay = nullptr; return bm::fail(BM_ERR_NOMEM, "layer table"); }
	bm::LayerTab t;
	std::memset(&t, 0, sizeof(t));
	for (int l = 0; l < kMaxStackLayers; l++) {
		t.A[l] = o->lay_set[0] ? o->lay[0][l] : (float)o->p.A;
		t.alpha[l] = o->lay_set[1] ? o->lay[1][l] : (float)o->p.alpha;
		t.K1[l] = o->lay_set[2] ? o->lay[2][l] : (float)o->p.K1;
		t.grp[l] = o->lay_grp[l];
		t.ihz2[l] = o->lay_ihz2[l];
		for (int c = 0; c < 3; c++) { t.selfc[l][c] = o->lay_selfc[l][c]; t.Ha[l][c] = o->lay_Ha[l][c]; }
	}
	BM_HIP(hipMemcpy(o->d_lay, &t, sizeof(t), hipMemcpyHostToDevice));
	return BM_OK;
}

// Self-demag coefficients of the evaluation speedup: the mesh's (DemagTFunc::SelfDemag_PBC
// with the mesh's periodic images, cell size in metres, Boris/DemagCUDA.cpp:37-39) and, for a
// layer stack, each layer mesh's own (SDemagCUDA_Demag::selfDemagCoeff).
int set_self_coeffs(bm_ode_s* o)
{
	double sd[3];
	if (o->pbc[0] || o->pbc[1] || o->pbc[2])
		bm::self_demag_coeff_pbc(o->h[0], o->h[1], o->h[2], o->nx, o->ny, o->nz, o->pbc[0], o->pbc[1], o->pbc[2], sd);
	else bm::self_demag_coeff(o->h[0], o->h[1], o->h[2], sd);
	for (int c = 0; c < 3; c++) o->selfc[c] = (float)sd[c];
	if (o->ml)
		for (int l = 0; l < o->nz && l < kMaxStackLayers; l++) {
			double hl[3];
			bm::mlayer_cell(o->ml, l, hl);
			bm::self_demag_coeff(hl[0], hl[1], hl[2], sd);
			for (int c = 0; c < 3; c++) o->lay_selfc[l][c] = (float)sd[c];
		}
	return upload_layers(o);
}

// Demag evaluations go through the fused row kernel (single open-boundary mesh, no speedup;
// periodic meshes take the unfused path, whose stage kernel wraps exchange around).
bool row_path(const bm_ode_s* o)
{
	return (o->p.modules & BM_MOD_DEMAG) && o->use_row && !o->ml && !o->speedup && !o->pbc[0] && !o->pbc[1] && !o->pbc[2]
		&& o->demag && !o->demag->full[0] && o->demag->Nx == 2 * o->nx && o->nx <= 2048;
}

// One field evaluation + stage `step` (the effective field is that of Mcur at stagetime).
int evaluate(bm_ode_s* o, int step, float dT, hipStream_t st)
{
	o->evaluations++;
	// the other modules' UpdateField of this stage (Boris/MeshBaseModules.cpp:98-111)
	if (o->ext_fn) o->ext_fn(o->ext_user, o->Mcur, o->Hext, (void*)st);
	const bool demag = (o->p.modules & BM_MOD_DEMAG) != 0;
	if (demag && o->speedup) return eval_speedup(o, step, dT, st);
	if (demag && o->ml) { int rc = bm::mlayer_planar(o->ml, o->Mcur, o->Hd, st); if (rc) return rc; o->convolutions++; return launch_stage(o, step, dT, st); }
	if (row_path(o)) { o->convolutions++; return eval_row(o, step, dT, st); }
	if (demag) {
		if (!o->Hd) return fail(BM_ERR_STATE, "unfused demag evaluation without its field buffer");
		int rc = bm::demag_planar(o->demag, o->Mcur, o->Hd, st); if (rc) return rc; o->convolutions++;
	}
	return launch_stage(o, step, dT, st);
}

// ODECommon_Base::SetAdaptiveTimeStep (Boris/DiffEq_CommonBase_Iterate.cpp:15-59).
bool set_adaptive_time_step(bm_ode_s* o)
{
	if (o->dT_min == o->dT_max) return true;
	if (o->lte > o->err_high_fail && o->dT > o->dT_min) {
		o->time -= o->dT; o->stagetime -= o->dT; o->iteration--; o->stageiteration--;
		double c = std::pow(o->err_high_fail * 0.8 / o->lte, 1.0 / (o->order + 1));
		if (c < 0.01) c = 0.01;
		o->dT *= c;
		if (o->dT < o->dT_min) o->dT = o->dT_min;
		o->available = false;
		o->rejections++;
		return false;
	}
	else if (o->lte > 0) {
		double c = std::pow(o->err_high_fail * 0.8 / o->lte, 1.0 / (o->order + 1));
		if (c > o->dT_increase) c = o->dT_increase;
		if (c < 0.01) c = 0.01;
		o->dT *= c;
		if (o->dT < o->dT_min) o->dT = o->dT_min;
		if (o->dT > o->dT_max) o->dT = o->dT_max;
	}
	return true;
}

int read_lte(bm_ode_s* o, hipStream_t st)
{
	BM_HIP(hipMemcpyAsync(o->h_lte, o->d_lte, sizeof(unsigned), hipMemcpyDeviceToHost, st));
	BM_HIP(hipStreamSynchronize(st));
	float f;
	std::memcpy(&f, o->h_lte, sizeof(float));
	o->lte = (double)f;
	return BM_OK;
}

int restore(bm_ode_s* o, hipStream_t st)
{
	BM_HIP(hipMemcpyAsync(o->M, o->sM1, o->n * 3 * sizeof(float), hipMemcpyDeviceToDevice, st));
	o->Mcur = o->M;
	o->S_valid = false;
	return BM_OK;
}

void end_step(bm_ode_s* o)
{
	o->evalStep = 0; o->time += o->dT; o->stagetime += o->dT; o->iteration++; o->stageiteration++; o->available = true;
}

// One field update + one stage: ODECommon_Base::IterateCUDA.
int iterate_stage(bm_ode_s* o, hipStream_t st)
{
	int rc;
	o->dT_last = o->dT;
	switch (o->mcode) {
	case M_RKDP: {
		if (o->evalStep == 0) {
			BM_HIP(hipMemsetAsync(o->d_lte, 0, sizeof(unsigned), st));
			if ((rc = evaluate(o, 0, (float)o->dT, st))) return rc;
			o->available = false;
			if (o->primed) {
				o->dT_last = o->dT;
				if ((rc = read_lte(o, st))) return rc;
				if (!set_adaptive_time_step(o)) {
					if ((rc = restore(o, st))) return rc;
					o->primed = false;
					break;
				}
			}
			else o->primed = true;
			if (row_path(o)) {
				// the advance rides on the x-R2C pass of the next evaluation's input
				if ((rc = bm::demag_fwd_x_advance(o->demag, o->Mcur, o->E[0], o->sM1, o->Mb, (float)o->dT / 5, st))) return rc;
				o->S_valid = true;
			}
			else {
				bm::prof_mark_begin(bm::PK_ADVANCE, st);
				hipLaunchKernelGGL(k_rkdp_advance, dim3(stage_grid(o)), dim3(kBS), 0, st, o->Mcur, o->Mb, o->sM1, o->E[0], o->n * 3, (float)o->dT / 5);
				bm::prof_mark_end(bm::PK_ADVANCE, st);
			}
			o->Mcur = o->Mb;
			o->evalStep++;
		}
		else if (o->evalStep < 5) {
			if ((rc = evaluate(o, o->evalStep, (float)o->dT, st))) return rc;
			o->evalStep++;
		}
		else {
			if ((rc = evaluate(o, 5, (float)o->dT, st))) return rc;
			end_step(o);
		}
	} break;
	case M_RK4: {
		if ((rc = evaluate(o, o->evalStep, (float)o->dT, st))) return rc;
		if (o->evalStep < 3) { o->evalStep++; o->available = false; }
		else end_step(o);
	} break;
	case M_TEULER:
	case M_AHEUN: {
		if (o->evalStep == 0) {
			if ((rc = evaluate(o, 0, (float)o->dT, st))) return rc;
			o->evalStep = 1; o->available = false;
		}
		else {
			if (o->mcode == M_AHEUN) BM_HIP(hipMemsetAsync(o->d_lte, 0, sizeof(unsigned), st));
			if ((rc = evaluate(o, 1, (float)o->dT, st))) return rc;
			end_step(o);
			if (o->mcode == M_AHEUN) {
				o->dT_last = o->dT;
				if ((rc = read_lte(o, st))) return rc;
				if (!set_adaptive_time_step(o)) { if ((rc = restore(o, st))) return rc; }
			}
		}
	} break;
	}
	BM_HIP(hipGetLastError());
	return BM_OK;
}

int advance_time(bm_ode_s* o, hipStream_t st)
{
	int rc;
	do { if ((rc = iterate_stage(o, st))) return rc; } while (!o->available);
	return BM_OK;
}

void free_all(bm_ode_s* o)
{
	if (o->Hd) (void)hipFree(o->Hd);
	if (o->sM1) (void)hipFree(o->sM1);
	if (o->Mb) (void)hipFree(o->Mb);
	for (int q = 0; q < 6; q++) if (o->E[q]) (void)hipFree(o->E[q]);
	if (o->Heff_keep) (void)hipFree(o->Heff_keep);
	for (int q = 0; q < 6; q++) if (o->Hs[q]) (void)hipFree(o->Hs[q]);
	if (o->d_lte) (void)hipFree(o->d_lte);
	if (o->h_lte) (void)hipHostFree(o->h_lte);
	if (o->d_part) (void)hipFree(o->d_part);
	if (o->h_part) (void)hipHostFree(o->h_part);
	if (o->Ms_layers) (void)hipFree(o->Ms_layers);
	if (o->d_lay) (void)hipFree(o->d_lay);
}

} // namespace

using bm::fail;

extern "C" int bm_ode_create(int method, int nx, int ny, int nz, double hx, double hy, double hz,
	const bm_fm_params* p, bm_demag_t demag, float* M, bm_ode_t* out)
{
	if (!out || !p || !M || nx < 1 || ny < 1 || nz < 1) return fail(BM_ERR_ARG, "bm_ode_create: bad arguments");
	*out = nullptr;
	if ((p->modules & BM_MOD_DEMAG) && (!demag || demag->nx != nx || demag->ny != ny || demag->nz != nz))
		return fail(BM_ERR_ARG, "bm_ode_create: demag handle missing or of another mesh size");
	bm_ode_s* o = new bm_ode_s();
	o->method = method; o->nx = nx; o->ny = ny; o->nz = nz; o->n = (size_t)nx * ny * nz;
	o->h[0] = hx; o->h[1] = hy; o->h[2] = hz; o->p = *p; o->demag = demag; o->M = M;
	o->ncu = bm::device_cu_count();
	// BM_ROW_KERNEL=0 selects the unfused evaluation (planar H_demag + stage kernel), kept
	// as a cross-check of the fused row kernel
	if (const char* e = getenv("BM_ROW_KERNEL")) o->use_row = atoi(e) != 0;
	// the row kernel's x passes need rows of exactly N.x / 2 cells with a zero-padded half: not
	// for a demag handle periodic along x, nor for rows shorter than N.x / 2 (general sizes),
	// nor for rows beyond the row kernels' 2048 cells (those take the unfused evaluation, whose
	// planar demag field buffer is allocated below)
	if (demag && (demag->full[0] || demag->Nx != 2 * nx || nx > 2048)) o->use_row = false;
	int nE = 0;
	// method defaults: Boris/DiffEq_CommonBase_Control.cpp:10-140 with Boris/DiffEq_Defs.h
	switch (method) {
	case BM_EVAL_RKDP54: o->mcode = M_RKDP; nE = 6; o->dT = 0.5e-12; o->err_high_fail = 1e-4; o->dT_increase = 2; o->dT_max = 3e-12; o->dT_min = 1e-15; o->order = 5; break;
	case BM_EVAL_RK4: o->mcode = M_RK4; nE = 3; o->dT = 0.5e-12; o->order = 4; break;
	case BM_EVAL_TEULER: o->mcode = M_TEULER; o->dT = 0.5e-13; o->order = 2; break;
	case BM_EVAL_AHEUN: o->mcode = M_AHEUN; o->dT = 0.5e-13; o->err_high_fail = 1e-4; o->dT_increase = 2; o->dT_max = 1e-12; o->dT_min = 1e-15; o->order = 2; break;
	default: delete o; return fail(BM_ERR_ARG, "bm_ode_create: unsupported evaluation method");
	}
	o->dT_last = o->dT;
	const size_t b3 = o->n * 3 * sizeof(float);
	bool ok = hipMalloc(&o->sM1, b3) == hipSuccess && hipMalloc(&o->Mb, b3) == hipSuccess;
	o->Mcur = o->M;
	for (int q = 0; q < nE && ok; q++) ok = hipMalloc(&o->E[q], b3) == hipSuccess && hipMemset(o->E[q], 0, b3) == hipSuccess;
	if (ok && (p->modules & BM_MOD_DEMAG) && !o->use_row) ok = hipMalloc(&o->Hd, b3) == hipSuccess;
	if (ok) ok = hipMalloc(&o->Heff_keep, b3) == hipSuccess;
	if (ok) ok = hipMalloc(&o->d_lte, sizeof(unsigned)) == hipSuccess;
	if (ok) ok = hipHostMalloc(&o->h_lte, sizeof(unsigned)) == hipSuccess;
	o->npart = (int)std::max<size_t>(1, std::min<size_t>((o->n + kBS - 1) / kBS, 1024));
	if (ok) ok = hipMalloc(&o->d_part, (size_t)o->npart * 3 * sizeof(double)) == hipSuccess;
	if (ok) ok = hipHostMalloc(&o->h_part, (size_t)o->npart * 3 * sizeof(double)) == hipSuccess;
	if (ok) ok = hipMemset(o->sM1, 0, b3) == hipSuccess;
	if (!ok) { free_all(o); delete o; return fail(BM_ERR_NOMEM, "bm_ode_create: device allocation failed"); }
	*out = o;
	return BM_OK;
}

extern "C" int bm_ode_destroy(bm_ode_t o)
{
	if (!o) return BM_OK;
	(void)hipDeviceSynchronize();
	free_all(o);
	delete o;
	return BM_OK;
}

extern "C" int bm_ode_set_ms_map(bm_ode_t o, const float* Ms_map)
{
	if (!o) return fail(BM_ERR_ARG, "null handle");
	o->Ms_map = Ms_map;
	return BM_OK;
}

// Layer stack: the ODE's nz planes are the layers of the multilayered demag ml.  Boris holds
// such a stack as one ferromagnetic mesh per layer coupled by SDemag (Boris/SDemag.cpp:173-300,
// SDemagCUDA::UpdateField); each mesh's exchange, anisotropy and LLG use that mesh's own
// parameters (Boris/Exchange.cpp:108, Boris/Anisotropy.cpp:82, Boris/DiffEqFM_Equations.cpp:12),
// so here exchange acts within each layer only and A, alpha, K1 (and Ms, through a per-cell
// map built here) may differ per layer.
extern "C" int bm_ode_set_layers(bm_ode_t o, bm_mlayer_t ml, const double* Ms, const double* A, const double* alpha, const double* K1)
{
	if (!o) return fail(BM_ERR_ARG, "null handle");
	if (!ml) {
		if ((o->p.modules & BM_MOD_DEMAG) && !o->demag) return fail(BM_ERR_STATE, "bm_ode_set_layers: demag is on and there is no other demag handle");
		if (o->Ms_map == o->Ms_layers) o->Ms_map = nullptr;
		o->ml = nullptr;
		o->num_saved = 0;
		return BM_OK;
	}
	int L, nx, ny;
	bm::mlayer_dims(ml, &L, &nx, &ny, nullptr);
	if (bm::mlayer_mapped(ml)) return fail(BM_ERR_SIZE, "bm_ode_set_layers: the layers must share one planned grid");
	if (L != o->nz || nx != o->nx || ny != o->ny) return fail(BM_ERR_ARG, "bm_ode_set_layers: the stack's layers must be the mesh's nz planes (same nx, ny)");
	int mp[2];
	bm::mlayer_pbc(ml, mp);
	if ((o->pbc[0] || o->pbc[1] || o->pbc[2]) && (o->pbc[2] || mp[0] != o->pbc[0] || mp[1] != o->pbc[1]))
		return fail(BM_ERR_STATE, "bm_ode_set_layers: the mesh's periodic images differ from the multilayer handle's");
	if ((mp[0] || mp[1]) && o->speedup) return fail(BM_ERR_STATE, "bm_ode_set_layers: no evaluation speedup on a periodic stack");
	if (L > bm::kMaxStackLayers) return fail(BM_ERR_SIZE, "bm_ode_set_layers: at most 64 layers");
	const size_t b3 = o->n * 3 * sizeof(float);
	if (!o->Hd && hipMalloc(&o->Hd, b3) != hipSuccess) return fail(BM_ERR_NOMEM, "bm_ode_set_layers: demag field buffer");
	const double* src[3] = { A, alpha, K1 };
	for (int q = 0; q < 3; q++) {
		o->lay_set[q] = src[q] != nullptr;
		for (int l = 0; l < bm::kMaxStackLayers; l++) o->lay[q][l] = src[q] && l < L ? (float)src[q][l] : 0.f;
	}
	if (Ms) {
		for (int l = 0; l < L; l++) if (!(Ms[l] > 0)) return fail(BM_ERR_ARG, "bm_ode_set_layers: Ms must be positive");
		if (!o->Ms_layers && hipMalloc(&o->Ms_layers, o->n * sizeof(float)) != hipSuccess) return fail(BM_ERR_NOMEM, "bm_ode_set_layers: Ms map");
		std::vector<float> h(o->n);
		const size_t nl = (size_t)nx * ny;
		for (size_t i = 0; i < o->n; i++) h[i] = (float)Ms[i / nl];
		BM_HIP(hipMemcpy(o->Ms_layers, h.data(), o->n * sizeof(float), hipMemcpyHostToDevice));
		o->Ms_map = o->Ms_layers;
	}
	o->ml = ml;
	o->use_row = false;
	o->S_valid = false;
	o->num_saved = 0;              // saved demag fields belong to the previous configuration
	o->lay_grouped = o->lay_field = false;
	return set_self_coeffs(o);
}

// Stacks of 3D meshes: consecutive layers of the stack with the same mesh index are the z
// planes of one ferromagnetic mesh of Boris's multilayered set-up (each mesh's own exchange
// couples its planes, Boris/Exchange.cpp:108 with BorisLib/VEC_VC_Del.h:10-130: free ends at
// the mesh's top and bottom, spacing = the planes' thickness; no exchange between meshes).
// Ha (3 doubles per layer, or null) gives each layer a constant applied field, as each mesh's
// own Zeeman module does (Boris/Zeeman.cpp:164).  The planes of a mesh must be touching and
// of one thickness.
extern "C" int bm_ode_set_layer_meshes(bm_ode_t o, const int* mesh_of_layer, const double* Ha)
{
	if (!o) return fail(BM_ERR_ARG, "null handle");
	if (!o->ml) return fail(BM_ERR_STATE, "bm_ode_set_layer_meshes: set a layer stack first (bm_ode_set_layers)");
	const int L = o->nz;
	if (Ha && o->p.wave != 0) return fail(BM_ERR_STATE, "bm_ode_set_layer_meshes: per-layer fields are constant (no applied-field waveform)");
	if (mesh_of_layer) {
		for (int l = 1; l < L; l++) {
			if (mesh_of_layer[l] < mesh_of_layer[l - 1]) return fail(BM_ERR_ARG, "bm_ode_set_layer_meshes: the planes of a mesh must be consecutive layers");
			if (mesh_of_layer[l] != mesh_of_layer[l - 1]) continue;
			double h0[3], h1[3];
			bm::mlayer_cell(o->ml, l - 1, h0); bm::mlayer_cell(o->ml, l, h1);
			const double gap = bm::mlayer_z0(o->ml, l) - bm::mlayer_z0(o->ml, l - 1);
			if (std::fabs(h0[2] - h1[2]) > 1e-6 * h0[2] || std::fabs(gap - h0[2]) > 1e-6 * h0[2])
				return fail(BM_ERR_ARG, "bm_ode_set_layer_meshes: planes of one mesh must be touching and of equal thickness");
		}
	}
	for (int l = 0; l < L; l++) {
		double hl[3];
		bm::mlayer_cell(o->ml, l, hl);
		o->lay_grp[l] = mesh_of_layer ? mesh_of_layer[l] : l;
		o->lay_ihz2[l] = (float)(1.0 / (hl[2] * hl[2]));
		for (int c = 0; c < 3; c++) o->lay_Ha[l][c] = Ha ? (float)Ha[3 * l + c] : 0.f;
	}
	o->lay_grouped = mesh_of_layer != nullptr;
	o->lay_field = Ha != nullptr;
	o->num_saved = 0;
	return upload_layers(o);
}

extern "C" int bm_ode_set_params(bm_ode_t o, const bm_fm_params* p)
{
	if (!o || !p) return fail(BM_ERR_ARG, "null argument");
	if ((p->modules & BM_MOD_DEMAG) && !o->ml && (!o->demag || o->demag->nx != o->nx || o->demag->ny != o->ny || o->demag->nz != o->nz))
		return fail(BM_ERR_STATE, "bm_ode_set_params: demag needs a demag handle of this mesh size at creation, or a layer stack");
	// the same checks bm_ode_set_layer_meshes and bm_ode_set_pbc make, for settings that
	// reach this call in the other order
	if (o->lay_field && p->wave != 0)
		return fail(BM_ERR_STATE, "bm_ode_set_params: per-layer constant fields are set (bm_ode_set_layer_meshes); no applied-field waveform");
	if ((p->modules & BM_MOD_DEMAG) && !o->ml && (o->pbc[0] || o->pbc[1] || o->pbc[2])
		&& (o->demag->pbc[0] != o->pbc[0] || o->demag->pbc[1] != o->pbc[1] || o->demag->pbc[2] != o->pbc[2]))
		return fail(BM_ERR_STATE, "bm_ode_set_params: the mesh is periodic and the demag handle has other periodic images");
	if ((p->modules & BM_MOD_DEMAG) && (!o->use_row || o->speedup) && !o->Hd) {
		const size_t b3 = o->n * 3 * sizeof(float);
		if (hipMalloc(&o->Hd, b3) != hipSuccess) return fail(BM_ERR_NOMEM, "bm_ode_set_params: demag field buffer");
	}
	o->p = *p;
	o->S_valid = false;
	o->num_saved = 0;      // DemagCUDA::UpdateConfiguration drops the saved fields
	return upload_layers(o);   // a stack's per-layer table holds the mesh values it defaults to
}

// Periodic boundary conditions (the `pbc` command: Demag::Set_PBC and Mesh::Set_Magnetic_PBC,
// Boris/Demag.cpp:137-148, Boris/MeshControl.cpp:49-65, Boris/SuperMeshMeshes_Settings.cpp:
// 659-700): exchange takes the cell on the other side as the neighbour at the mesh ends of
// every axis with images, and with demag on the demag handle must have been created with the
// same images.  Saved speedup fields are dropped and the speedup's self coefficient gains the
// images' far-field terms.
extern "C" int bm_ode_set_pbc(bm_ode_t o, int pbc_x, int pbc_y, int pbc_z)
{
	if (!o || pbc_x < 0 || pbc_y < 0 || pbc_z < 0) return fail(BM_ERR_ARG, "bm_ode_set_pbc: bad arguments");
	if (o->ml) {
		// a layer stack: the multilayer handle carries the x, y images (bm_mlayer_create_pbc)
		int mp[2];
		bm::mlayer_pbc(o->ml, mp);
		if (pbc_z || mp[0] != pbc_x || mp[1] != pbc_y)
			return fail(BM_ERR_STATE, "bm_ode_set_pbc: a layer stack takes the x, y images of its multilayer handle (bm_mlayer_create_pbc), none along z");
		if (o->speedup && (pbc_x || pbc_y)) return fail(BM_ERR_STATE, "bm_ode_set_pbc: no evaluation speedup on a periodic stack");
	}
	else if ((o->p.modules & BM_MOD_DEMAG) && o->demag
		&& (o->demag->pbc[0] != pbc_x || o->demag->pbc[1] != pbc_y || o->demag->pbc[2] != pbc_z))
		return fail(BM_ERR_STATE, "bm_ode_set_pbc: the demag handle must have the same periodic images (bm_demag_create_pbc)");
	o->pbc[0] = pbc_x; o->pbc[1] = pbc_y; o->pbc[2] = pbc_z;
	if ((o->p.modules & BM_MOD_DEMAG) && !row_path(o) && !o->Hd) {
		if (hipMalloc(&o->Hd, o->n * 3 * sizeof(float)) != hipSuccess) { o->Hd = nullptr; return fail(BM_ERR_NOMEM, "bm_ode_set_pbc: demag field buffer"); }
	}
	o->S_valid = false;
	o->num_saved = 0;
	return set_self_coeffs(o);
}

extern "C" int bm_ode_set_eval_speedup(bm_ode_t o, int mode, double dTspeedup)
{
	if (!o || mode < 0 || mode > 6) return fail(BM_ERR_ARG, "bm_ode_set_eval_speedup: mode must be 0..6");
	if (mode && o->ml && (o->pbc[0] || o->pbc[1])) return fail(BM_ERR_STATE, "bm_ode_set_eval_speedup: no evaluation speedup on a periodic stack");
	const size_t b3 = o->n * 3 * sizeof(float);
	for (int q = 0; q < 6; q++) {
		if (q >= mode && o->Hs[q]) { (void)hipFree(o->Hs[q]); o->Hs[q] = nullptr; }
		if (q < mode && !o->Hs[q] && hipMalloc(&o->Hs[q], b3) != hipSuccess) {
			o->Hs[q] = nullptr;
			return fail(BM_ERR_NOMEM, "bm_ode_set_eval_speedup: saved field buffers");
		}
	}
	if (mode && !o->Hd && hipMalloc(&o->Hd, b3) != hipSuccess) { o->Hd = nullptr; return fail(BM_ERR_NOMEM, "bm_ode_set_eval_speedup: demag field buffer"); }
	o->speedup = mode;
	o->num_saved = 0;
	for (int q = 0; q < 6; q++) o->tdm[q] = 0;
	// DemagTFunc::SelfDemag_PBC without images, cell size in metres (Boris/DemagCUDA.cpp:37)
	if (int rc = set_self_coeffs(o)) return rc;
	if (dTspeedup > 0) { o->link_dTspeedup = false; o->dTspeedup = dTspeedup; o->time_speedup = o->time; }
	else { o->link_dTspeedup = true; o->dTspeedup = o->dT; }
	o->S_valid = false;
	return BM_OK;
}

extern "C" int bm_ode_get_speedup_state(bm_ode_t o, int* num_saved, double* tdm6, long long* convolutions)
{
	if (!o) return fail(BM_ERR_ARG, "null handle");
	if (num_saved) *num_saved = o->num_saved;
	if (tdm6) for (int q = 0; q < 6; q++) tdm6[q] = o->tdm[q];
	if (convolutions) *convolutions = o->convolutions;
	return BM_OK;
}

// Fields of modules outside the fused four (Boris/MeshBaseModules.cpp:98-111: every module's
// UpdateField runs before each stage; here the caller's function computes their sum from the
// stage's magnetisation, and the stage kernel adds it).
extern "C" int bm_ode_set_extra_field(bm_ode_t o, bm_field_fn fn, void* user, float* Hext)
{
	if (!o) return fail(BM_ERR_ARG, "null handle");
	if (fn && !Hext) return fail(BM_ERR_ARG, "bm_ode_set_extra_field: a field function needs its output array");
	o->ext_fn = fn;
	o->ext_user = fn ? user : nullptr;
	o->Hext = fn ? Hext : nullptr;
	return BM_OK;
}

extern "C" int bm_ode_set_dt(bm_ode_t o, double dT)
{
	if (!o || !(dT > 0)) return fail(BM_ERR_ARG, "bad dT");
	o->dT = dT; o->dT_last = dT;
	return BM_OK;
}

extern "C" int bm_ode_set_astep(bm_ode_t o, double err, double inc, double dtmin, double dtmax)
{
	if (!o) return fail(BM_ERR_ARG, "null handle");
	if (err > 0) o->err_high_fail = err;
	if (inc > 0) o->dT_increase = inc;
	if (dtmin > 0) o->dT_min = dtmin;
	if (dtmax > 0) o->dT_max = dtmax;
	return BM_OK;
}

extern "C" int bm_ode_advance_time(bm_ode_t o, void* stream)
{
	if (!o) return fail(BM_ERR_ARG, "null handle");
	o->S_valid = false;    // the caller may have written M since the last call
	return advance_time(o, (hipStream_t)stream);
}

extern "C" int bm_ode_run_to_iteration(bm_ode_t o, long long iteration, void* stream)
{
	if (!o) return fail(BM_ERR_ARG, "null handle");
	o->S_valid = false;
	while (o->iteration < iteration) { int rc = advance_time(o, (hipStream_t)stream); if (rc) return rc; }
	return BM_OK;
}

extern "C" int bm_ode_iterate_stage(bm_ode_t o, void* stream)
{
	if (!o) return fail(BM_ERR_ARG, "null handle");
	o->S_valid = false;
	return iterate_stage(o, (hipStream_t)stream);
}

extern "C" int bm_ode_restore(bm_ode_t o, void* stream)
{
	if (!o) return fail(BM_ERR_ARG, "null handle");
	return restore(o, (hipStream_t)stream);
}

extern "C" int bm_ode_get_state(bm_ode_t o, bm_ode_state* s)
{
	if (!o || !s) return fail(BM_ERR_ARG, "null argument");
	s->time = o->time; s->stagetime = o->stagetime; s->dT = o->dT; s->dT_last = o->dT_last; s->lte = o->lte;
	s->iteration = o->iteration; s->stageiteration = o->stageiteration; s->rejections = o->rejections; s->evaluations = o->evaluations;
	s->evalStep = o->evalStep; s->primed = o->primed; s->available = o->available;
	return BM_OK;
}

extern "C" int bm_ode_keep_heff(bm_ode_t o, int on)
{
	if (!o) return fail(BM_ERR_ARG, "null handle");
	o->keep_heff = on != 0;
	return BM_OK;
}

extern "C" int bm_ode_copy_heff(bm_ode_t o, float* H, void* stream)
{
	if (!o || !H) return fail(BM_ERR_ARG, "null argument");
	if (!o->keep_heff) return fail(BM_ERR_STATE, "bm_ode_copy_heff: enable bm_ode_keep_heff first");
	BM_HIP(hipMemcpyAsync(H, o->Heff_keep, o->n * 3 * sizeof(float), hipMemcpyDeviceToDevice, (hipStream_t)stream));
	return BM_OK;
}

extern "C" int bm_ode_average_m(bm_ode_t o, double* out3, void* stream)
{
	if (!o || !out3) return fail(BM_ERR_ARG, "null argument");
	hipStream_t st = (hipStream_t)stream;
	hipLaunchKernelGGL(k_avg_partial, dim3(o->npart), dim3(kBS), 0, st, o->Mcur, o->Ms_map, (float)o->p.Ms, o->n, o->d_part);
	BM_HIP(hipMemcpyAsync(o->h_part, o->d_part, (size_t)o->npart * 3 * sizeof(double), hipMemcpyDeviceToHost, st));
	BM_HIP(hipStreamSynchronize(st));
	double s[3] = { 0, 0, 0 };
	for (int b = 0; b < o->npart; b++) for (int c = 0; c < 3; c++) s[c] += o->h_part[3 * b + c];
	for (int c = 0; c < 3; c++) out3[c] = s[c] / (double)o->n;
	return BM_OK;
}
