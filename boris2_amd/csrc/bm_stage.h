// Per-cell effective field + LLG + explicit-stage update, shared by the stand-alone stage
// kernel (demag field read from a planar buffer, one cell per thread) and the fused row
// kernel (demag field produced in registers by the x-direction inverse FFT, two adjacent
// cells of a row per call).
//
// Effective field in the reference's module order — Zeeman sets it (Boris/ZeemanCUDA.cu:31),
// demag adds, exchange adds (6-neighbour Laplacian with free-boundary one-sided differences,
// Boris/ExchangeCUDA.cu:15 + BorisLib/VEC_VC_Del.h:10-130), uniaxial anisotropy adds
// (Boris/AnisotropyUniCUDA.cu:13) — then the LLG right-hand side
// (Boris/DiffEqFM_Equations.cpp:12-24) and the stage update of the chosen method:
//   RKDP54  Boris/DiffEqFM_EvalsCUDA_RKDP54.cu:15-330  (FSAL lte at stage 0)
//   RK4     Boris/DiffEqFM_EvalsLLGCUDA.cu:416-641
//   TEuler  Boris/DiffEqFM_EvalsLLGCUDA.cu:121-255 (Heun)
//   AHeun   Boris/DiffEqFM_EvalsLLGCUDA.cu:299-363
#pragma once

#include <hip/hip_runtime.h>
#include "boris_mi355x.h"
#include "bm_fft.h"

namespace bm {

constexpr int kBS = 256;
constexpr double kMU0 = 1.256637061435916e-6;    // BorisLib/Funcs_Math_base.h:9
constexpr double kGAMMA = 2.212761569e5;         // BorisLib/Funcs_Math_base.h:16

struct F3 { float x, y, z; };
__device__ __forceinline__ F3 ld3(const float* p, size_t i) { return F3{ p[3 * i], p[3 * i + 1], p[3 * i + 2] }; }
__device__ __forceinline__ void st3(float* p, size_t i, F3 v) { p[3 * i] = v.x; p[3 * i + 1] = v.y; p[3 * i + 2] = v.z; }
__device__ __forceinline__ F3 operator+(F3 a, F3 b) { return F3{ a.x + b.x, a.y + b.y, a.z + b.z }; }
__device__ __forceinline__ F3 operator-(F3 a, F3 b) { return F3{ a.x - b.x, a.y - b.y, a.z - b.z }; }
__device__ __forceinline__ F3 operator*(float s, F3 a) { return F3{ s * a.x, s * a.y, s * a.z }; }
__device__ __forceinline__ F3 cross(F3 a, F3 b) { return F3{ a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x }; }
__device__ __forceinline__ float dot(F3 a, F3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ float norm(F3 a) { return sqrtf(dot(a, a)); }

// Cells 2k and 2k+1 of a row: their 6 floats are contiguous in every AoS array, so they
// move as three 8-byte accesses (idx even, arrays 8-byte aligned).
struct P6 { F3 a, b; };
__device__ __forceinline__ P6 operator+(P6 u, P6 v) { return P6{ u.a + v.a, u.b + v.b }; }
__device__ __forceinline__ P6 operator-(P6 u, P6 v) { return P6{ u.a - v.a, u.b - v.b }; }
__device__ __forceinline__ P6 operator*(float s, P6 u) { return P6{ s * u.a, s * u.b }; }
__device__ __forceinline__ P6 ld6(const float* p, size_t i)
{
	const float2* q = reinterpret_cast<const float2*>(p + 3 * i);
	const float2 u = q[0], v = q[1], w = q[2];
	return P6{ F3{ u.x, u.y, v.x }, F3{ v.y, w.x, w.y } };
}
__device__ __forceinline__ void st6(float* p, size_t i, P6 x)
{
	float2* q = reinterpret_cast<float2*>(p + 3 * i);
	q[0] = make_float2(x.a.x, x.a.y); q[1] = make_float2(x.a.z, x.b.x); q[2] = make_float2(x.b.y, x.b.z);
}

template <class V> __device__ __forceinline__ V ldv(const float* p, size_t i);
template <> __device__ __forceinline__ F3 ldv<F3>(const float* p, size_t i) { return ld3(p, i); }
template <> __device__ __forceinline__ P6 ldv<P6>(const float* p, size_t i) { return ld6(p, i); }
__device__ __forceinline__ void stv(float* p, size_t i, F3 v) { st3(p, i, v); }
__device__ __forceinline__ void stv(float* p, size_t i, P6 v) { st6(p, i, v); }
// renormalisation to the cell's Ms and the relative difference used as lte
__device__ __forceinline__ F3 renorm(F3 v, float Ms) { return (Ms / norm(v)) * v; }
__device__ __forceinline__ P6 renorm(P6 v, float2 Ms) { return P6{ renorm(v.a, Ms.x), renorm(v.b, Ms.y) }; }
__device__ __forceinline__ float rel_err(F3 m, F3 p) { return norm(m - p) / norm(m); }
__device__ __forceinline__ float rel_err(P6 m, P6 p) { return fmaxf(rel_err(m.a, p.a), rel_err(m.b, p.b)); }

constexpr int kMaxStackLayers = 64;
// a batched layer stack keeps every layer's spectra in registers in its kernel product
constexpr int kMaxBatchStackLayers = 16;

// Per-layer values of a layer stack, kept in device memory (one table per stack, written at
// configuration) and indexed by the cell's layer, so the stage arguments stay small.
struct LayerTab {
	float A[kMaxStackLayers], alpha[kMaxStackLayers], K1[kMaxStackLayers];
	float selfc[kMaxStackLayers][3];   // evaluation speedup: each layer's own self coefficient
	// stacks of 3D meshes: layers with equal grp are consecutive z planes of one mesh, coupled
	// by exchange across planes with 1 / hz^2 = ihz2; Ha: each layer's constant applied field
	int grp[kMaxStackLayers];
	float ihz2[kMaxStackLayers];
	float Ha[kMaxStackLayers][3];
};

// Per-mesh adaptive step control of a batched RKDP54 integration, kept on the device so a
// captured step graph runs it (ODECommon_Base's time, dT, lte, iteration, primed state and
// SetAdaptiveTimeStep controls, Boris/DiffEq_CommonBase.h, Boris/DiffEq_CommonBase_Iterate.cpp:15-59).
// `time` comes first: the stage kernels read it as the step start time of time-dependent fields.
struct BatchCtl {
	double time, dT, lte;
	double err_high_fail, dT_increase, dT_min, dT_max;
	long long iteration, target, rejections;
	int order, primed;
	int active;   // the mesh still has accepted steps to make in this run
	int skip;     // this attempt was rejected (or the mesh is inactive): no stages 1..5
};

struct StageArgs {
	const float* Min;       // magnetisation this stage evaluates (neighbours read)
	float* Mout;            // magnetisation of the next stage (never the same array)
	const float* Hd;        // planar demag field (3 planes of n) or null (stand-alone kernel)
	float* sM1;
	float* E[6];
	const float* Ms_map;
	float* Heff_out;        // optional AoS copy of the effective field
	const float* Hext;      // optional AoS field of other modules, added to the effective field
	unsigned* lte;          // max lte as float bits
	int nx, ny, nz;
	int per;                // periodic exchange along x, y, z: bits 0, 1, 2
	float ihx2, ihy2, ihz2; // 1 / h^2 per axis
	float Ms, A, alpha, K1, K2, grel;
	float ea[3], Ha[3];
	float dT;
	int modules;
	int map;
	// layer stack: the nz planes are separate 2D layers (one ferromagnetic mesh each, no
	// exchange between them) with their own A, alpha and K1 (when layered != 0)
	int layered;
	const LayerTab* lay;    // per-layer values (device memory)
	// stacks of 3D meshes: exchange across the planes of a mesh (lay_grouped != 0) and
	// per-layer applied fields (lay_field != 0), from lay
	int lay_grouped, lay_field;
	// time-dependent applied field read on the device (batched graphs): with tnow set, the
	// stage kernel evaluates Ha0 + amp dir sinc(w (t - t0)) (wave 1) or Ha0 + slope t (wave 2)
	// at t = *tnow in FP64 and rounds it to FP32, as the host does for a single mesh; a
	// one-thread-per-mesh kernel does that once per step into Ha_dev (3 floats), which the
	// stage kernels then read instead of Ha
	const double* tnow;
	const float* Ha_dev;
	int wave;
	double Ha0[3], amp, w, t0, dir[3], slope[3];
	// demag evaluation speedup (Boris/DemagCUDA.cpp:106-468, kernels Boris/DemagCUDA.cu:44-413):
	// with nsp > 0 the demag field of a cell is sum_{s<nsp} ca[s] Hs[s] + selfc * M (planar
	// saved fields, AoS M of this stage); with Hsave set, the stage also stores Hd - selfc * M
	// (planar) after a full convolution.
	const float* Hs[6];
	float* Hsave;
	float ca[6], selfc[3];
	int nsp;
	// batched adaptive integration: the mesh's step control; stage 0 runs while the mesh is
	// active, stages 1..5 unless the attempt was rejected
	const BatchCtl* ctl;
};

// Applied field of a stage: Ha, or the mesh's device-held value when Ha_dev is set.
__device__ __forceinline__ F3 stage_Ha(const StageArgs& a)
{
	if (a.Ha_dev) return F3{ a.Ha_dev[0], a.Ha_dev[1], a.Ha_dev[2] };
	return F3{ a.Ha[0], a.Ha[1], a.Ha[2] };
}

// The time-dependent applied field at the device-held step time *a.tnow (sinc pulse or ramp
// over the constant part), in FP64 rounded to FP32 (see StageArgs::tnow).
__device__ __forceinline__ F3 waveform_Ha(const StageArgs& a)
{
	const double t = *a.tnow;
	double H[3] = { a.Ha0[0], a.Ha0[1], a.Ha0[2] };
	if (a.wave == 1) {
		const double x = a.w * (t - a.t0);
		const double sc = (fabs(x) < 1e-300) ? 1.0 : sin(x) / x;
		for (int c = 0; c < 3; c++) H[c] += a.amp * a.dir[c] * sc;
	}
	else if (a.wave == 2) for (int c = 0; c < 3; c++) H[c] += a.slope[c] * t;
	return F3{ (float)H[0], (float)H[1], (float)H[2] };
}

// Material constants of a cell: the mesh's, or those of its layer in a layer stack (the
// layer value is picked with compile-time indices so StageArgs stays in registers).
struct CellMat { float A, alpha, K1; };
__device__ __forceinline__ CellMat cell_mat(const StageArgs& a, int k)
{
	if (!a.layered) return CellMat{ a.A, a.alpha, a.K1 };
	return CellMat{ a.lay->A[k], a.lay->alpha[k], a.lay->K1[k] };
}

enum { M_RKDP = 0, M_RK4 = 1, M_TEULER = 2, M_AHEUN = 3 };

// Block-wide max of a non-negative float, folded into *dst with an integer atomicMax on
// its bit pattern (order-independent, so deterministic).  All NT threads of the block must
// call it.
template <int NT = kBS>
__device__ __forceinline__ void block_max_to(float v, unsigned* dst)
{
	__shared__ float red[NT / 64];
	for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o));
	if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
	__syncthreads();
	if (threadIdx.x == 0) {
		float m = red[0];
		for (int q = 1; q < NT / 64; q++) m = fmaxf(m, red[q]);
		atomicMax(dst, __float_as_uint(m));
	}
}

// Effective field of one cell; lap = sum over axes of the second differences / h^2.
__device__ __forceinline__ F3 eff_field(const StageArgs& a, const CellMat& cm, F3 Ha, F3 m, F3 Hd, F3 lap, float Ms)
{
	F3 H = (a.modules & BM_MOD_ZEEMAN) ? Ha : F3{ 0.f, 0.f, 0.f };
	if (a.modules & BM_MOD_DEMAG) H = H + Hd;
	if (a.modules & BM_MOD_EXCHANGE) H = H + ((2.f * cm.A) / ((float)kMU0 * Ms * Ms)) * lap;
	if (a.modules & BM_MOD_ANIS_UNI) {
		const F3 e{ a.ea[0], a.ea[1], a.ea[2] };
		const float d = dot(m, e) / Ms;
		H = H + ((2.f / ((float)kMU0 * Ms)) * d * (cm.K1 + 2.f * a.K2 * (1.f - d * d))) * e;
	}
	return H;
}

// LLG right-hand side (Landau-Lifshitz form with the Gilbert damping folded in)
__device__ __forceinline__ F3 llg_rhs(const StageArgs& a, float alpha, F3 m, F3 H, float Ms)
{
	const F3 mxh = cross(m, H);
	const float iMs = 1.f / Ms;
	const F3 mu{ m.x * iMs, m.y * iMs, m.z * iMs };
	const float pre = (float)(-kGAMMA) * a.grel / (1.f + alpha * alpha);
	return pre * (mxh + alpha * cross(mu, mxh));
}

// Stage update of one cell (V = F3, MS = float) or of a cell pair (V = P6, MS = float2):
// stores E_s, sM1 and the output magnetisation as the method requires and returns the
// output magnetisation (m itself for RKDP54 stage 0, which writes none).
template <int METHOD, int STEP, class V, class MS>
__device__ __forceinline__ V rk_update(const StageArgs& a, size_t idx, V m, V rhs, MS Ms, float& lte_loc)
{
	const float dT = a.dT;
	V mn = m;
	if (METHOD == M_RKDP) {
		if (STEP == 0) {
			// embedded 4th-order solution: stage 5 left all of it but the b*_7 k_7 term in E1
			const V pred = ldv<V>(a.E[1], idx) + (dT * (1.f / 40.f)) * rhs;
			lte_loc = fmaxf(lte_loc, rel_err(m, pred));
			stv(a.E[0], idx, rhs);
			return m;
		}
		else if (STEP == 1) {
			stv(a.E[1], idx, rhs);
			mn = ldv<V>(a.sM1, idx) + dT * ((3.f / 40.f) * ldv<V>(a.E[0], idx) + (9.f / 40.f) * rhs);
		}
		else if (STEP == 2) {
			stv(a.E[2], idx, rhs);
			mn = ldv<V>(a.sM1, idx) + dT * ((44.f / 45.f) * ldv<V>(a.E[0], idx) - (56.f / 15.f) * ldv<V>(a.E[1], idx) + (32.f / 9.f) * rhs);
		}
		else if (STEP == 3) {
			stv(a.E[3], idx, rhs);
			mn = ldv<V>(a.sM1, idx) + dT * ((19372.f / 6561.f) * ldv<V>(a.E[0], idx) - (25360.f / 2187.f) * ldv<V>(a.E[1], idx)
				+ (64448.f / 6561.f) * ldv<V>(a.E[2], idx) - (212.f / 729.f) * rhs);
		}
		else if (STEP == 4) {
			stv(a.E[4], idx, rhs);
			mn = ldv<V>(a.sM1, idx) + dT * ((9017.f / 3168.f) * ldv<V>(a.E[0], idx) - (355.f / 33.f) * ldv<V>(a.E[1], idx)
				+ (46732.f / 5247.f) * ldv<V>(a.E[2], idx) + (49.f / 176.f) * ldv<V>(a.E[3], idx) - (5103.f / 18656.f) * rhs);
		}
		else {
			stv(a.E[5], idx, rhs);
			const V s1 = ldv<V>(a.sM1, idx), e0 = ldv<V>(a.E[0], idx), e2 = ldv<V>(a.E[2], idx), e3 = ldv<V>(a.E[3], idx), e4 = ldv<V>(a.E[4], idx);
			mn = s1 + dT * ((35.f / 384.f) * e0 + (500.f / 1113.f) * e2 + (125.f / 192.f) * e3 - (2187.f / 6784.f) * e4 + (11.f / 84.f) * rhs);
			// the embedded 4th-order solution except its k_7 term (added by the next stage 0,
			// which evaluates k_7 = f(new M)) goes to E1, free after stage 4: stage 0 then
			// reads one array instead of six
			stv(a.E[1], idx, s1 + dT * ((5179.f / 57600.f) * e0 + (7571.f / 16695.f) * e2 + (393.f / 640.f) * e3
				- (92097.f / 339200.f) * e4 + (187.f / 2100.f) * rhs));
			mn = renorm(mn, Ms);
		}
	}
	else if (METHOD == M_RK4) {
		if (STEP == 0) { stv(a.sM1, idx, m); stv(a.E[0], idx, rhs); mn = m + (dT / 2) * rhs; }
		else if (STEP == 1) { stv(a.E[1], idx, rhs); mn = ldv<V>(a.sM1, idx) + (dT / 2) * rhs; }
		else if (STEP == 2) { stv(a.E[2], idx, rhs); mn = ldv<V>(a.sM1, idx) + dT * rhs; }
		else {
			mn = ldv<V>(a.sM1, idx) + (dT / 6) * (ldv<V>(a.E[0], idx) + 2.f * ldv<V>(a.E[1], idx) + 2.f * ldv<V>(a.E[2], idx) + rhs);
			mn = renorm(mn, Ms);
		}
	}
	else {
		if (STEP == 0) { stv(a.sM1, idx, m); mn = m + dT * rhs; }
		else {
			mn = 0.5f * (ldv<V>(a.sM1, idx) + m + dT * rhs);
			mn = renorm(mn, Ms);
			if (METHOD == M_AHEUN) lte_loc = fmaxf(lte_loc, rel_err(mn, m));
		}
	}
	stv(a.Mout, idx, mn);
	return mn;
}

// One cell of the stand-alone stage kernel (any mesh size).
template <int METHOD, int STEP>
__device__ __forceinline__ void stage_cell(const StageArgs& a, F3 Ha, size_t idx, int i, int j, int k, F3 Hd, float& lte_loc)
{
	const F3 m = ld3(a.Min, idx);
	const float Ms = a.Ms_map ? a.Ms_map[idx] : a.Ms;
	F3 lap{ 0.f, 0.f, 0.f };
	if (a.modules & BM_MOD_EXCHANGE) {
		F3 lx{ 0, 0, 0 }, ly{ 0, 0, 0 }, lz{ 0, 0, 0 };
		const size_t sy = a.nx, sz = (size_t)a.nx * a.ny;
		// periodic axes at the mesh ends: (M[+] + M[-]) - 2M with the wrapped neighbour
		// (BorisLib/VEC_VC_Del.h:24-45)
		if (a.nx > 1) {
			if (i > 0 && i < a.nx - 1) lx = ld3(a.Min, idx + 1) + ld3(a.Min, idx - 1) - 2.f * m;
			else if (a.per & 1) lx = (i == 0) ? ld3(a.Min, idx + 1) + ld3(a.Min, idx + (a.nx - 1)) - 2.f * m
			                                  : ld3(a.Min, idx - (a.nx - 1)) + ld3(a.Min, idx - 1) - 2.f * m;
			else if (i == 0) lx = ld3(a.Min, idx + 1) - m;
			else lx = ld3(a.Min, idx - 1) - m;
		}
		if (a.ny > 1) {
			if (j > 0 && j < a.ny - 1) ly = ld3(a.Min, idx + sy) + ld3(a.Min, idx - sy) - 2.f * m;
			else if (a.per & 2) ly = (j == 0) ? ld3(a.Min, idx + sy) + ld3(a.Min, idx + (a.ny - 1) * sy) - 2.f * m
			                                  : ld3(a.Min, idx - (a.ny - 1) * sy) + ld3(a.Min, idx - sy) - 2.f * m;
			else if (j == 0) ly = ld3(a.Min, idx + sy) - m;
			else ly = ld3(a.Min, idx - sy) - m;
		}
		float ihz2 = a.ihz2;
		if (a.nz > 1 && !a.layered) {
			if (k > 0 && k < a.nz - 1) lz = ld3(a.Min, idx + sz) + ld3(a.Min, idx - sz) - 2.f * m;
			else if (a.per & 4) lz = (k == 0) ? ld3(a.Min, idx + sz) + ld3(a.Min, idx + (a.nz - 1) * sz) - 2.f * m
			                                  : ld3(a.Min, idx - (a.nz - 1) * sz) + ld3(a.Min, idx - sz) - 2.f * m;
			else if (k == 0) lz = ld3(a.Min, idx + sz) - m;
			else lz = ld3(a.Min, idx - sz) - m;
		}
		else if (a.lay_grouped && a.nz > 1) {
			// planes of one 3D mesh in a stack: neighbours only within the mesh (free ends)
			const int* grp = a.lay->grp;
			const int g = grp[k];
			const bool lo = k > 0 && grp[k - 1] == g, hi = k < a.nz - 1 && grp[k + 1] == g;
			if (lo && hi) lz = ld3(a.Min, idx + sz) + ld3(a.Min, idx - sz) - 2.f * m;
			else if (hi) lz = ld3(a.Min, idx + sz) - m;
			else if (lo) lz = ld3(a.Min, idx - sz) - m;
			ihz2 = a.lay->ihz2[k];
		}
		lap = a.ihx2 * lx + a.ihy2 * ly + ihz2 * lz;
	}
	const CellMat cm = cell_mat(a, k);
	if (a.lay_field) Ha = F3{ a.lay->Ha[k][0], a.lay->Ha[k][1], a.lay->Ha[k][2] };
	F3 H = eff_field(a, cm, Ha, m, Hd, lap, Ms);
	if (a.Hext) H = H + ld3(a.Hext, idx);
	if (a.Heff_out) st3(a.Heff_out, idx, H);
	(void)rk_update<METHOD, STEP>(a, idx, m, llg_rhs(a, cm.alpha, m, H, Ms), Ms, lte_loc);
}

// Cells i0 = 2k and i0 + 1 of row (j, k) for the row kernel (nx even, so both exist).  Row
// and plane boundaries are uniform over a row, so only the x ends take a per-lane select.
// Returns the output magnetisation of both cells.
template <int METHOD, int STEP>
__device__ __forceinline__ P6 stage_pair(const StageArgs& a, size_t idx, int i0, int j, int k, P6 m, P6 Hd, float& lte_loc)
{
	const float2 Ms = a.Ms_map ? *reinterpret_cast<const float2*>(a.Ms_map + idx) : make_float2(a.Ms, a.Ms);
	F3 lap0{ 0.f, 0.f, 0.f }, lap1{ 0.f, 0.f, 0.f };
	if (a.modules & BM_MOD_EXCHANGE) {
		// x: the two cells are each other's neighbours; clamped loads keep the ends branch-free
		// (the row kernel serves open-boundary meshes only)
		const bool hasL = i0 > 0, hasR = i0 + 2 < a.nx;
		const F3 L = ld3(a.Min, hasL ? idx - 1 : idx), R = ld3(a.Min, hasR ? idx + 2 : idx + 1);
		F3 lx0 = hasL ? (m.b + L) - 2.f * m.a : m.b - m.a;
		F3 lx1 = hasR ? (R + m.a) - 2.f * m.b : m.a - m.b;
		P6 ly{ F3{ 0, 0, 0 }, F3{ 0, 0, 0 } }, lz = ly;
		const size_t sy = a.nx, sz = (size_t)a.nx * a.ny;
		if (a.ny > 1) {
			if (j > 0 && j < a.ny - 1) ly = ld6(a.Min, idx + sy) + ld6(a.Min, idx - sy) - 2.f * m;
			else if (j == 0) ly = ld6(a.Min, idx + sy) - m;
			else ly = ld6(a.Min, idx - sy) - m;
		}
		if (a.nz > 1) {
			if (k > 0 && k < a.nz - 1) lz = ld6(a.Min, idx + sz) + ld6(a.Min, idx - sz) - 2.f * m;
			else if (k == 0) lz = ld6(a.Min, idx + sz) - m;
			else lz = ld6(a.Min, idx - sz) - m;
		}
		lap0 = a.ihx2 * lx0 + a.ihy2 * ly.a + a.ihz2 * lz.a;
		lap1 = a.ihx2 * lx1 + a.ihy2 * ly.b + a.ihz2 * lz.b;
	}
	// the row kernel serves single meshes only (never a layer stack)
	const CellMat cm{ a.A, a.alpha, a.K1 };
	const F3 Ha{ a.Ha[0], a.Ha[1], a.Ha[2] };
	F3 H0 = eff_field(a, cm, Ha, m.a, Hd.a, lap0, Ms.x), H1 = eff_field(a, cm, Ha, m.b, Hd.b, lap1, Ms.y);
	if (a.Hext) { const P6 he = ld6(a.Hext, idx); H0 = H0 + he.a; H1 = H1 + he.b; }
	if (a.Heff_out) st6(a.Heff_out, idx, P6{ H0, H1 });
	const P6 rhs{ llg_rhs(a, cm.alpha, m.a, H0, Ms.x), llg_rhs(a, cm.alpha, m.b, H1, Ms.y) };
	return rk_update<METHOD, STEP>(a, idx, m, rhs, Ms, lte_loc);
}

// Batched stage kernel over nb meshes (fixed-step methods), bm_ode.hip.
int batch_stage(int mcode, int step, const StageArgs* d_tab, int nb, const int* d_choff, int nchunks, hipStream_t st);

// Stages that produce an lte (reduced over the mesh to a max)
template <int METHOD, int STEP> constexpr bool stage_has_lte() { return (METHOD == M_RKDP && STEP == 0) || (METHOD == M_AHEUN && STEP == 1); }

} // namespace bm
