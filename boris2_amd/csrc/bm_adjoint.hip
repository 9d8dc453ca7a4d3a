// Discrete adjoint of the fixed-step RK4 LLG rollout with demag, exchange, uniaxial
// anisotropy and a Zeeman field H(t) = Ha + amp * dir * sinc(w (t - t0)) or Ha + slope * t,
// per-cell Ms map; optionally a layer stack (the nz planes are 2D layers coupled by the
// multilayered demag, per-layer A, alpha, K1, exchange within layers).
//
// No reference counterpart (the reference has no differentiation); the forward scheme is
// the plain path's RK4 step (Boris/DiffEqFM_EvalsLLGCUDA.cu:416-641: Zeeman at the step's
// start time, renormalisation to the cell's Ms after the step).
//
// Loss L = 1 - (1/n) sum_i (M_i / Ms_i) . mt_i of the final state, M0_i = Ms_i m0_i.
// Gradients dL/dMs_i, dL/d amp and dL/d slope.  The adjoint of one RK4 step runs the stages
// backwards; each stage's vector-Jacobian product needs the field at the stage input (taped
// during the forward recomputation) and one demag convolution of the field adjoint u — the
// single-mesh demag operator is symmetric (N_ij = N_ji, N(r) = N(-r)), so its adjoint is the
// same convolution; the multilayered one runs its exact transpose (pair kernels swapped,
// imaginary xz/yz parts negated); the exchange Laplacian is symmetric (free boundaries, and
// periodic axes, whose wrapped stencil is circulant), with the per-cell prefactor
// 2A/(mu0 Ms^2) moved to the input side; anisotropy is local.  A periodic demag kernel sums
// images at +-k L, which keeps N(r) = N(-r), so periodic meshes reuse the same convolution.
//
// Memory: the state is checkpointed every K steps; the backward pass recomputes one segment
// of K steps at a time with a tape of the four stage inputs, their demag fields and the
// pre-renormalisation sum (27 floats per cell per step), then sweeps it backwards.  Cost:
// forward 4 convolutions per step; backward 4 (recompute) + 4 (adjoint) per step.
#include "bm_internal.h"
#include "bm_stage.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

using namespace bm;

// layer stacks in the adjoint: per-layer values passed in the kernel arguments (8 layers)
constexpr int kMaxAdjLayers = 8;

namespace {

constexpr int kAB = 256;

struct AdjArgs {
	int nx, ny, nz;
	float ihx2, ihy2, ihz2;
	float A, alpha, grel;
	float K1, K2, ea[3];    // uniaxial anisotropy
	float Hz[3];            // applied field at the step's start time
	float dir[3], sc;       // pulse direction and sinc value at that time (d H / d amp = sc dir)
	float tz;               // ramp: d H / d slope_c = tz e_c (the step's start time), else 0
	int modules;
	const float* Ms;
	int layered;            // layer stack: per-layer A, alpha, K1, no exchange across layers
	int per;                // periodic exchange along x, y, z: bits 0, 1, 2
	float lay_A[kMaxAdjLayers], lay_alpha[kMaxAdjLayers], lay_K1[kMaxAdjLayers];
};

__device__ __forceinline__ CellMat adj_mat(const AdjArgs& a, int k)
{
	if (!a.layered) return CellMat{ a.A, a.alpha, a.K1 };
	CellMat c{ a.lay_A[0], a.lay_alpha[0], a.lay_K1[0] };
#pragma unroll
	for (int l = 1; l < kMaxAdjLayers; l++)
		if (k == l) c = CellMat{ a.lay_A[l], a.lay_alpha[l], a.lay_K1[l] };
	return c;
}

__device__ __forceinline__ F3 lap_at(const float* X, const AdjArgs& a, size_t idx, int i, int j, int k)
{
	const F3 m = ld3(X, idx);
	F3 lx{ 0, 0, 0 }, ly{ 0, 0, 0 }, lz{ 0, 0, 0 };
	const size_t sy = a.nx, sz = (size_t)a.nx * a.ny;
	// periodic axes at the mesh ends: the wrapped neighbour (BorisLib/VEC_VC_Del.h:24-45)
	if (a.nx > 1) {
		if (i > 0 && i < a.nx - 1) lx = ld3(X, idx + 1) + ld3(X, idx - 1) - 2.f * m;
		else if (a.per & 1) lx = (i == 0) ? ld3(X, idx + 1) + ld3(X, idx + (a.nx - 1)) - 2.f * m
		                                  : ld3(X, idx - (a.nx - 1)) + ld3(X, idx - 1) - 2.f * m;
		else if (i == 0) lx = ld3(X, idx + 1) - m;
		else lx = ld3(X, idx - 1) - m;
	}
	if (a.ny > 1) {
		if (j > 0 && j < a.ny - 1) ly = ld3(X, idx + sy) + ld3(X, idx - sy) - 2.f * m;
		else if (a.per & 2) ly = (j == 0) ? ld3(X, idx + sy) + ld3(X, idx + (a.ny - 1) * sy) - 2.f * m
		                                  : ld3(X, idx - (a.ny - 1) * sy) + ld3(X, idx - sy) - 2.f * m;
		else if (j == 0) ly = ld3(X, idx + sy) - m;
		else ly = ld3(X, idx - sy) - m;
	}
	if (a.nz > 1 && !a.layered) {
		if (k > 0 && k < a.nz - 1) lz = ld3(X, idx + sz) + ld3(X, idx - sz) - 2.f * m;
		else if (a.per & 4) lz = (k == 0) ? ld3(X, idx + sz) + ld3(X, idx + (a.nz - 1) * sz) - 2.f * m
		                                  : ld3(X, idx - (a.nz - 1) * sz) + ld3(X, idx - sz) - 2.f * m;
		else if (k == 0) lz = ld3(X, idx + sz) - m;
		else lz = ld3(X, idx - sz) - m;
	}
	return a.ihx2 * lx + a.ihy2 * ly + a.ihz2 * lz;
}

__device__ __forceinline__ float cex(float A, float Ms) { return (2.f * A) / ((float)kMU0 * Ms * Ms); }
__device__ __forceinline__ float pre_llg(const AdjArgs& a, float alpha) { return (float)(-kGAMMA) * a.grel / (1.f + alpha * alpha); }

// field at X: Zeeman + demag (planar Hd) + exchange + anisotropy; also returns the unscaled
// Laplacian
__device__ __forceinline__ F3 field_at(const AdjArgs& a, const CellMat& cm, const float* X, const float* Hd, size_t n, size_t idx,
	int i, int j, int k, float Ms, F3& lap)
{
	F3 H = (a.modules & BM_MOD_ZEEMAN) ? F3{ a.Hz[0], a.Hz[1], a.Hz[2] } : F3{ 0.f, 0.f, 0.f };
	if (a.modules & BM_MOD_DEMAG) H = H + F3{ Hd[idx], Hd[n + idx], Hd[2 * n + idx] };
	lap = F3{ 0.f, 0.f, 0.f };
	if (a.modules & BM_MOD_EXCHANGE) { lap = lap_at(X, a, idx, i, j, k); H = H + cex(cm.A, Ms) * lap; }
	if (a.modules & BM_MOD_ANIS_UNI) {
		const F3 e{ a.ea[0], a.ea[1], a.ea[2] };
		const float d = dot(ld3(X, idx), e) / Ms;
		H = H + ((2.f / ((float)kMU0 * Ms)) * d * (cm.K1 + 2.f * a.K2 * (1.f - d * d))) * e;
	}
	return H;
}

__device__ __forceinline__ void cell_ijk(const AdjArgs& a, size_t idx, int& i, int& j, int& k)
{
	i = (int)(idx % a.nx); j = (int)((idx / a.nx) % a.ny); k = (int)(idx / ((size_t)a.nx * a.ny));
}

__global__ void k_init(float* M, const float* Ms, const float* m0, size_t n)
{
	for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
		st3(M, i, Ms[i] * ld3(m0, i));
}

// Forward stage S of the step from Mn: k = f(X); S < 3: acc (=, +=) w_S k, Xn = Mn + c_S dt k;
// S == 3: Y = Mn + dt/6 (acc + k) (stored when Yout), Mnext = Ms Y / |Y|.
template <int S>
__global__ void __launch_bounds__(kAB) k_fwd(AdjArgs a, const float* __restrict__ X, const float* __restrict__ Hd, const float* __restrict__ Mn,
	float* __restrict__ acc, float* __restrict__ Xn, float* __restrict__ Yout, float* __restrict__ Mnext, float dt)
{
	const size_t n = (size_t)a.nx * a.ny * a.nz;
	for (size_t idx = blockIdx.x * (size_t)blockDim.x + threadIdx.x; idx < n; idx += (size_t)gridDim.x * blockDim.x) {
		int i, j, k; cell_ijk(a, idx, i, j, k);
		const float Ms = a.Ms[idx];
		const CellMat cm = adj_mat(a, k);
		F3 lap;
		const F3 m = ld3(X, idx);
		const F3 H = field_at(a, cm, X, Hd, n, idx, i, j, k, Ms, lap);
		const F3 mxh = cross(m, H);
		const float iMs = 1.f / Ms;
		const F3 mu{ m.x * iMs, m.y * iMs, m.z * iMs };
		const F3 f = pre_llg(a, cm.alpha) * (mxh + cm.alpha * cross(mu, mxh));
		if (S == 0) { st3(acc, idx, f); st3(Xn, idx, ld3(Mn, idx) + (dt / 2) * f); }
		else if (S == 1) { st3(acc, idx, ld3(acc, idx) + 2.f * f); st3(Xn, idx, ld3(Mn, idx) + (dt / 2) * f); }
		else if (S == 2) { st3(acc, idx, ld3(acc, idx) + 2.f * f); st3(Xn, idx, ld3(Mn, idx) + dt * f); }
		else {
			const F3 Y = ld3(Mn, idx) + (dt / 6) * (ld3(acc, idx) + f);
			if (Yout) st3(Yout, idx, Y);
			st3(Mnext, idx, (Ms / norm(Y)) * Y);
		}
	}
}

// deterministic per-block partial sums of sum_i (M_i . mt_i) / Ms_i
__global__ void __launch_bounds__(kAB) k_loss_part(const float* M, const float* mt, const float* Ms, size_t n, double* part)
{
	__shared__ double red[kAB];
	double s = 0;
	for (size_t i = blockIdx.x * (size_t)kAB + threadIdx.x; i < n; i += (size_t)gridDim.x * kAB)
		s += (double)dot(ld3(M, i), ld3(mt, i)) / (double)Ms[i];
	red[threadIdx.x] = s;
	__syncthreads();
	for (int o = kAB / 2; o > 0; o >>= 1) { if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o]; __syncthreads(); }
	if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// lam = dL/dM of the final state, gMs = explicit Ms term of the loss
__global__ void k_final(const float* M, const float* mt, const float* Ms, size_t n, float* lam, double* gMs)
{
	for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
		const float ms = Ms[i];
		const F3 t = ld3(mt, i);
		st3(lam, i, (-1.f / ((float)n * ms)) * t);
		gMs[i] = (double)dot(ld3(M, i), t) / ((double)n * (double)ms * (double)ms);
	}
}

// renormalisation adjoint: lamY = (Ms/|Y|)(lam - y (y.lam)), gMs += y.lam; lacc = lamY
__global__ void k_adj_renorm(const float* Y, const float* Ms, const float* lam, float* lamY, float* lacc, double* gMs, size_t n)
{
	for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
		const F3 y = ld3(Y, i);
		const float nr = norm(y);
		const F3 yh = (1.f / nr) * y;
		const F3 l = ld3(lam, i);
		const float yl = dot(yh, l);
		gMs[i] += (double)yl;
		const F3 ly = (Ms[i] / nr) * (l - yl * yh);
		st3(lamY, i, ly);
		st3(lacc, i, ly);
	}
}

// Stage VJP, part 1: w = (dt/6) wS lamY + cS dt lxn; lx = direct term (anisotropy's local
// field Jacobian included); u (field adjoint) and c u (AoS); gMs += Ms terms; per-block
// partials of sum_i u_i . dir sc (amplitude) and sum_i u_i tz (slope), 4 per block.
__global__ void __launch_bounds__(kAB) k_vjp_a(AdjArgs a, const float* __restrict__ X, const float* __restrict__ Hd,
	const float* __restrict__ lamY, const float* __restrict__ lxn, float wS, float cS, float dt,
	float* __restrict__ lx, float* __restrict__ u, float* __restrict__ cu, double* __restrict__ gMs, double* __restrict__ apart)
{
	__shared__ double red[4][kAB];
	const size_t n = (size_t)a.nx * a.ny * a.nz;
	double gz[4] = { 0, 0, 0, 0 };
	for (size_t idx = blockIdx.x * (size_t)kAB + threadIdx.x; idx < n; idx += (size_t)gridDim.x * kAB) {
		int i, j, k; cell_ijk(a, idx, i, j, k);
		const float Ms = a.Ms[idx];
		const CellMat cm = adj_mat(a, k);
		const float pr = pre_llg(a, cm.alpha), al = cm.alpha;
		F3 lap;
		const F3 m = ld3(X, idx);
		const F3 H = field_at(a, cm, X, Hd, n, idx, i, j, k, Ms, lap);
		F3 w = ((dt / 6) * wS) * ld3(lamY, idx);
		if (lxn) w = w + (cS * dt) * ld3(lxn, idx);
		const float mh = dot(m, H), mw = dot(m, w), hw = dot(H, w), mm = dot(m, m);
		F3 d = pr * (cross(H, w) + (al / Ms) * (mh * w + mw * H - (2.f * hw) * m));
		const F3 uu = pr * (cross(w, m) + (al / Ms) * (mw * m - mm * w));
		const float ce = cex(cm.A, Ms);
		double g = -(double)pr * (double)al / ((double)Ms * Ms) * (double)dot(cross(m, cross(m, H)), w);
		if (a.modules & BM_MOD_EXCHANGE) g += (double)dot(uu, lap) * (-2.0 * (double)ce / (double)Ms);
		if (a.modules & BM_MOD_ANIS_UNI) {
			// H_an = 2/mu0 s (K1 + 2K2 - 2K2 s^2/Ms^2)/Ms^2 e with s = m.e: its Jacobian is
			// 2/(mu0 Ms^2) (K1 + 2K2 - 6K2 d^2) e e^T (d = s/Ms), its Ms derivative
			// 2/mu0 s e (-2K1 - 4K2 + 8K2 d^2)/Ms^3
			const F3 e{ a.ea[0], a.ea[1], a.ea[2] };
			const float sm = dot(m, e), dd = sm / Ms, ue = dot(uu, e);
			d = d + ((2.f / ((float)kMU0 * Ms * Ms)) * (cm.K1 + 2.f * a.K2 - 6.f * a.K2 * dd * dd) * ue) * e;
			g += 2.0 / kMU0 * (double)sm * (double)ue * (-2.0 * cm.K1 - 4.0 * a.K2 + 8.0 * a.K2 * (double)dd * dd) / ((double)Ms * Ms * Ms);
		}
		st3(lx, idx, d);
		st3(u, idx, uu);
		st3(cu, idx, ce * uu);
		gMs[idx] += g;
		gz[0] += (double)(dot(uu, F3{ a.dir[0], a.dir[1], a.dir[2] }) * a.sc);
		gz[1] += (double)(uu.x * a.tz); gz[2] += (double)(uu.y * a.tz); gz[3] += (double)(uu.z * a.tz);
	}
	for (int c = 0; c < 4; c++) red[c][threadIdx.x] = gz[c];
	__syncthreads();
	for (int o = kAB / 2; o > 0; o >>= 1) {
		if (threadIdx.x < o) for (int c = 0; c < 4; c++) red[c][threadIdx.x] += red[c][threadIdx.x + o];
		__syncthreads();
	}
	if (threadIdx.x < 4) apart[4 * blockIdx.x + threadIdx.x] += red[threadIdx.x][0];
}

// Stage VJP, part 2: lx += N u (planar) + Lap(c u); lacc += lx
__global__ void __launch_bounds__(kAB) k_vjp_b(AdjArgs a, const float* __restrict__ Nu, const float* __restrict__ cu,
	float* __restrict__ lx, float* __restrict__ lacc)
{
	const size_t n = (size_t)a.nx * a.ny * a.nz;
	for (size_t idx = blockIdx.x * (size_t)kAB + threadIdx.x; idx < n; idx += (size_t)gridDim.x * kAB) {
		int i, j, k; cell_ijk(a, idx, i, j, k);
		F3 v = ld3(lx, idx);
		if (a.modules & BM_MOD_DEMAG) v = v + F3{ Nu[idx], Nu[n + idx], Nu[2 * n + idx] };
		if (a.modules & BM_MOD_EXCHANGE) v = v + lap_at(cu, a, idx, i, j, k);
		st3(lx, idx, v);
		st3(lacc, idx, ld3(lacc, idx) + v);
	}
}

__global__ void k_m0_term(const float* m0, const float* lam, double* gMs, size_t n)
{
	for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
		gMs[i] += (double)dot(ld3(m0, i), ld3(lam, i));
}

__global__ void k_to_float(const double* g, float* out, size_t n)
{
	for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) out[i] = (float)g[i];
}

} // namespace

struct bm_adj_s {
	int nx, ny, nz; size_t n;
	double h[3];
	bm_fm_params p;
	bm_demag_t demag;
	bm_mlayer_t ml = nullptr;     // layer stack (bm_adj_set_layers)
	int pbc[3] = { 0, 0, 0 };     // periodic images per axis (bm_adj_set_pbc)
	bool lay_set[3] = { false, false, false };
	float lay[3][kMaxAdjLayers] = {};
	double gz[4] = { 0, 0, 0, 0 };  // dL/d amp, dL/d slope of the last backward pass
	const float* Ms;
	double dt;
	int nsteps, K, nck;
	int grid = 1024;
	// device buffers
	float* ck = nullptr;          // checkpoints: nck x 3n
	float* tape = nullptr;        // K steps x (X0..X3, Y) x 3n, then K x 4 demag fields x 3n
	float* Mw = nullptr;          // work states: 2 x 3n
	float* acc = nullptr;
	float* Hd = nullptr;          // planar 3n (forward without tape)
	float* Xs = nullptr;          // stage inputs without tape: 3 x 3n
	float* lam = nullptr; float* lamY = nullptr; float* lacc = nullptr; float* lx = nullptr; float* lxn = nullptr;
	float* u = nullptr; float* cu = nullptr; float* Nu = nullptr;
	double* gMs = nullptr;
	double* part = nullptr;       // per-block partials: loss (grid), then 4 Zeeman sums per block
	double* h_part = nullptr;     // pinned
	const float* m0 = nullptr; const float* mt = nullptr;
	bool forward_done = false;
	float* cur = nullptr;         // forward: the state the next step starts from
	// batched runs (bm_adj_batch_*): this handle's own stream and its completion event
	hipStream_t own = nullptr;
	hipEvent_t done = nullptr;
};

namespace {

AdjArgs adj_args(bm_adj_s* o, int step)
{
	AdjArgs a;
	std::memset(&a, 0, sizeof(a));
	a.nx = o->nx; a.ny = o->ny; a.nz = o->nz;
	a.ihx2 = (float)(1.0 / (o->h[0] * o->h[0])); a.ihy2 = (float)(1.0 / (o->h[1] * o->h[1])); a.ihz2 = (float)(1.0 / (o->h[2] * o->h[2]));
	a.A = (float)o->p.A; a.alpha = (float)o->p.alpha; a.grel = (float)o->p.grel;
	a.K1 = (float)o->p.K1; a.K2 = (float)o->p.K2;
	for (int c = 0; c < 3; c++) a.ea[c] = (float)o->p.ea[c];
	if (o->ml) {
		a.layered = 1;
		a.ihz2 = 0.f;
		for (int l = 0; l < kMaxAdjLayers; l++) {
			a.lay_A[l] = o->lay_set[0] ? o->lay[0][l] : a.A;
			a.lay_alpha[l] = o->lay_set[1] ? o->lay[1][l] : a.alpha;
			a.lay_K1[l] = o->lay_set[2] ? o->lay[2][l] : a.K1;
		}
	}
	const double t = step * o->dt;
	double sc = 0.0;
	if (o->p.wave == 1) { const double x = o->p.w * (t - o->p.t0); sc = (std::fabs(x) < 1e-300) ? 1.0 : std::sin(x) / x; }
	for (int c = 0; c < 3; c++) {
		a.Hz[c] = (float)(o->p.Ha[c] + (o->p.wave == 1 ? o->p.amp * o->p.dir[c] * sc : 0.0) + (o->p.wave == 2 ? o->p.slope[c] * t : 0.0));
		a.dir[c] = (float)o->p.dir[c];
	}
	a.sc = (o->p.wave == 1 && (o->p.modules & BM_MOD_ZEEMAN)) ? (float)sc : 0.f;
	a.tz = (o->p.wave == 2 && (o->p.modules & BM_MOD_ZEEMAN)) ? (float)t : 0.f;
	a.modules = o->p.modules;
	a.Ms = o->Ms;
	a.per = (o->pbc[0] ? 1 : 0) | (o->pbc[1] ? 2 : 0) | (o->pbc[2] ? 4 : 0);
	return a;
}

size_t n3(bm_adj_s* o) { return o->n * 3; }
float* tX(bm_adj_s* o, int q, int s) { return o->tape + ((size_t)q * 5 + s) * n3(o); }                 // s = 0..3 stage inputs, 4 = Y
float* tH(bm_adj_s* o, int q, int s) { return o->tape + (size_t)o->K * 5 * n3(o) + ((size_t)q * 4 + s) * n3(o); }

// one forward step from Mn into Mnext; with taping (q >= 0) the stage inputs, their demag
// fields and Y go to tape slot q
int fwd_step(bm_adj_s* o, int step, const float* Mn, float* Mnext, int q, hipStream_t st)
{
	const AdjArgs a = adj_args(o, step);
	const float dt = (float)o->dt;
	const float* X = Mn;
	const dim3 g(o->grid), b(kAB);
	for (int s = 0; s < 4; s++) {
		float* Hd = q >= 0 ? tH(o, q, s) : o->Hd;
		if (q >= 0 && s == 0) BM_HIP(hipMemcpyAsync(tX(o, q, 0), Mn, n3(o) * sizeof(float), hipMemcpyDeviceToDevice, st));
		if (o->p.modules & BM_MOD_DEMAG) {
			int rc = o->ml ? mlayer_planar(o->ml, X, Hd, st) : demag_planar(o->demag, X, Hd, st);
			if (rc) return rc;
		}
		float* Xn = s < 3 ? (q >= 0 ? tX(o, q, s + 1) : o->Xs + (size_t)s * n3(o)) : nullptr;
		float* Y = (s == 3 && q >= 0) ? tX(o, q, 4) : nullptr;
		switch (s) {
		case 0: hipLaunchKernelGGL(k_fwd<0>, g, b, 0, st, a, X, Hd, Mn, o->acc, Xn, Y, Mnext, dt); break;
		case 1: hipLaunchKernelGGL(k_fwd<1>, g, b, 0, st, a, X, Hd, Mn, o->acc, Xn, Y, Mnext, dt); break;
		case 2: hipLaunchKernelGGL(k_fwd<2>, g, b, 0, st, a, X, Hd, Mn, o->acc, Xn, Y, Mnext, dt); break;
		default: hipLaunchKernelGGL(k_fwd<3>, g, b, 0, st, a, X, Hd, Mn, o->acc, Xn, Y, Mnext, dt); break;
		}
		X = Xn;
	}
	return BM_OK;
}

// adjoint of one step with tape slot q: lam (dL/dM_{step+1}) -> lam (dL/dM_step)
int adj_step(bm_adj_s* o, int step, int q, hipStream_t st)
{
	const AdjArgs a = adj_args(o, step);
	const float dt = (float)o->dt;
	const dim3 g(o->grid), b(kAB);
	static const float wgt[4] = { 1, 2, 2, 1 }, cst[3] = { 0.5f, 0.5f, 1.0f };
	hipLaunchKernelGGL(k_adj_renorm, g, b, 0, st, tX(o, q, 4), o->Ms, o->lam, o->lamY, o->lacc, o->gMs, o->n);
	for (int s = 3; s >= 0; s--) {
		const float* lxn = s < 3 ? o->lxn : nullptr;
		hipLaunchKernelGGL(k_vjp_a, g, b, 0, st, a, tX(o, q, s), tH(o, q, s), o->lamY, lxn, wgt[s], s < 3 ? cst[s] : 0.f, dt,
			o->lx, o->u, o->cu, o->gMs, o->part + o->grid);
		if (o->p.modules & BM_MOD_DEMAG) {
			int rc = o->ml ? mlayer_planar_transposed(o->ml, o->u, o->Nu, st) : demag_planar(o->demag, o->u, o->Nu, st);
			if (rc) return rc;
		}
		hipLaunchKernelGGL(k_vjp_b, g, b, 0, st, a, o->Nu, o->cu, o->lx, o->lacc);
		std::swap(o->lx, o->lxn);
	}
	std::swap(o->lam, o->lacc);
	return BM_OK;
}

void free_all(bm_adj_s* o)
{
	if (o->done) (void)hipEventDestroy(o->done);
	if (o->own) (void)hipStreamDestroy(o->own);
	float* fl[] = { o->ck, o->tape, o->Mw, o->acc, o->Hd, o->Xs, o->lam, o->lamY, o->lacc, o->lx, o->lxn, o->u, o->cu, o->Nu };
	for (float* p : fl) if (p) (void)hipFree(p);
	if (o->gMs) (void)hipFree(o->gMs);
	if (o->part) (void)hipFree(o->part);
	if (o->h_part) (void)hipHostFree(o->h_part);
}

} // namespace

extern "C" int bm_adj_create(int nx, int ny, int nz, double hx, double hy, double hz, const bm_fm_params* p,
	bm_demag_t demag, const float* Ms_map, double dt, int nsteps, int ckpt_every, bm_adj_t* out)
{
	if (!out || !p || !Ms_map || nx < 1 || ny < 1 || nz < 1 || !(dt > 0) || nsteps < 1 || ckpt_every < 1)
		return fail(BM_ERR_ARG, "bm_adj_create: bad arguments");
	*out = nullptr;
	if ((p->modules & BM_MOD_DEMAG) && (!demag || demag->nx != nx || demag->ny != ny || demag->nz != nz))
		return fail(BM_ERR_ARG, "bm_adj_create: demag handle missing or of another mesh size");
	bm_adj_s* o = new bm_adj_s();
	o->nx = nx; o->ny = ny; o->nz = nz; o->n = (size_t)nx * ny * nz;
	o->h[0] = hx; o->h[1] = hy; o->h[2] = hz; o->p = *p; o->demag = demag; o->Ms = Ms_map;
	o->dt = dt; o->nsteps = nsteps; o->K = std::min(ckpt_every, nsteps);
	o->nck = (nsteps + o->K - 1) / o->K + 1;
	o->grid = (int)std::max<size_t>(1, std::min<size_t>((o->n + kAB - 1) / kAB, (size_t)device_cu_count() * 8));
	const size_t b3 = n3(o) * sizeof(float);
	bool ok = hipMalloc(&o->ck, (size_t)o->nck * b3) == hipSuccess
		&& hipMalloc(&o->tape, (size_t)o->K * 9 * b3) == hipSuccess
		&& hipMalloc(&o->Mw, 2 * b3) == hipSuccess && hipMalloc(&o->acc, b3) == hipSuccess
		&& hipMalloc(&o->Hd, b3) == hipSuccess && hipMalloc(&o->Xs, 3 * b3) == hipSuccess
		&& hipMalloc(&o->lam, b3) == hipSuccess && hipMalloc(&o->lamY, b3) == hipSuccess && hipMalloc(&o->lacc, b3) == hipSuccess
		&& hipMalloc(&o->lx, b3) == hipSuccess && hipMalloc(&o->lxn, b3) == hipSuccess
		&& hipMalloc(&o->u, b3) == hipSuccess && hipMalloc(&o->cu, b3) == hipSuccess && hipMalloc(&o->Nu, b3) == hipSuccess
		&& hipMalloc(&o->gMs, o->n * sizeof(double)) == hipSuccess
		&& hipMalloc(&o->part, 5 * (size_t)o->grid * sizeof(double)) == hipSuccess
		&& hipHostMalloc(&o->h_part, 5 * (size_t)o->grid * sizeof(double)) == hipSuccess;
	if (!ok) { free_all(o); delete o; return fail(BM_ERR_NOMEM, "bm_adj_create: device memory"); }
	*out = o;
	return BM_OK;
}

extern "C" int bm_adj_destroy(bm_adj_t o)
{
	if (!o) return BM_OK;
	(void)hipDeviceSynchronize();
	free_all(o);
	delete o;
	return BM_OK;
}

namespace {

// Forward rollout in phases, so that several handles can be driven in lockstep.
int fwd_begin(bm_adj_s* o, const float* m0, const float* mt, hipStream_t st)
{
	o->m0 = m0; o->mt = mt;
	hipLaunchKernelGGL(k_init, dim3(o->grid), dim3(kAB), 0, st, o->ck, o->Ms, m0, o->n);
	o->cur = o->ck;
	return BM_OK;
}

// step s, checkpointing every K steps and at the end
int fwd_advance(bm_adj_s* o, int s, hipStream_t st)
{
	const bool at_ck = (s + 1) % o->K == 0 || s + 1 == o->nsteps;
	float* next = at_ck ? o->ck + (size_t)((s + 1 + o->K - 1) / o->K) * n3(o) : (o->cur == o->Mw ? o->Mw + n3(o) : o->Mw);
	int rc = fwd_step(o, s, o->cur, next, -1, st);
	o->cur = next;
	return rc;
}

int fwd_end(bm_adj_s* o, float* M_final, hipStream_t st)
{
	if (M_final) BM_HIP(hipMemcpyAsync(M_final, o->cur, n3(o) * sizeof(float), hipMemcpyDeviceToDevice, st));
	hipLaunchKernelGGL(k_loss_part, dim3(o->grid), dim3(kAB), 0, st, o->cur, o->mt, o->Ms, o->n, o->part);
	BM_HIP(hipMemcpyAsync(o->h_part, o->part, (size_t)o->grid * sizeof(double), hipMemcpyDeviceToHost, st));
	return BM_OK;
}

// after the stream has finished: the loss from the per-block partials (fixed order)
void fwd_finish(bm_adj_s* o, double* loss)
{
	double sum = 0;
	for (int q = 0; q < o->grid; q++) sum += o->h_part[q];
	if (loss) *loss = 1.0 - sum / (double)o->n;
	o->forward_done = true;
}

int bwd_begin(bm_adj_s* o, hipStream_t st)
{
	float* Mfin = o->ck + (size_t)(o->nck - 1) * n3(o);
	hipLaunchKernelGGL(k_final, dim3(o->grid), dim3(kAB), 0, st, Mfin, o->mt, o->Ms, o->n, o->lam, o->gMs);
	BM_HIP(hipMemsetAsync(o->part + o->grid, 0, 4 * (size_t)o->grid * sizeof(double), st));
	return BM_OK;
}

// segment sg: recompute its steps from the checkpoint with the tape, then run them backwards
int bwd_recompute(bm_adj_s* o, int sg, int s, hipStream_t st)
{
	const int s0 = sg * o->K;
	if (s == s0) o->cur = o->ck + (size_t)sg * n3(o);
	float* next = (o->cur == o->Mw) ? o->Mw + n3(o) : o->Mw;
	int rc = fwd_step(o, s, o->cur, next, s - s0, st);
	o->cur = next;
	return rc;
}

int bwd_end(bm_adj_s* o, float* gMs_out, hipStream_t st)
{
	const dim3 g(o->grid), b(kAB);
	hipLaunchKernelGGL(k_m0_term, g, b, 0, st, o->m0, o->lam, o->gMs, o->n);
	hipLaunchKernelGGL(k_to_float, g, b, 0, st, o->gMs, gMs_out, o->n);
	BM_HIP(hipMemcpyAsync(o->h_part, o->part + o->grid, 4 * (size_t)o->grid * sizeof(double), hipMemcpyDeviceToHost, st));
	return BM_OK;
}

void bwd_finish(bm_adj_s* o, double* gamp)
{
	for (int c = 0; c < 4; c++) o->gz[c] = 0;
	for (int q = 0; q < o->grid; q++) for (int c = 0; c < 4; c++) o->gz[c] += o->h_part[4 * q + c];
	if (gamp) *gamp = o->gz[0];
}

// The streams of a batched run: every handle's own stream starts after earlier work on the
// caller's stream; `join` makes the caller's stream wait for all of them.
int batch_fork(bm_adj_s* const* hs, int count, hipStream_t st)
{
	hipEvent_t ev;
	BM_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
	int rc = BM_OK;
	if (hipEventRecord(ev, st) != hipSuccess) rc = fail(BM_ERR_HIP, "bm_adj_batch: event");
	for (int b = 0; b < count && !rc; b++) {
		bm_adj_s* o = hs[b];
		if (!o->own && (hipStreamCreate(&o->own) != hipSuccess || hipEventCreateWithFlags(&o->done, hipEventDisableTiming) != hipSuccess))
			rc = fail(BM_ERR_HIP, "bm_adj_batch: stream");
		else if (hipStreamWaitEvent(o->own, ev, 0) != hipSuccess) rc = fail(BM_ERR_HIP, "bm_adj_batch: wait");
	}
	(void)hipEventDestroy(ev);
	return rc;
}

int batch_join(bm_adj_s* const* hs, int count, hipStream_t st)
{
	for (int b = 0; b < count; b++) {
		BM_HIP(hipEventRecord(hs[b]->done, hs[b]->own));
		BM_HIP(hipStreamWaitEvent(st, hs[b]->done, 0));
	}
	BM_HIP(hipStreamSynchronize(st));
	return BM_OK;
}

int batch_check(int count, bm_adj_s* const* hs)
{
	if (count < 1 || !hs) return fail(BM_ERR_ARG, "bm_adj_batch: no handles");
	for (int b = 0; b < count; b++) {
		if (!hs[b]) return fail(BM_ERR_ARG, "bm_adj_batch: null handle");
		if (hs[b]->nsteps != hs[0]->nsteps || hs[b]->K != hs[0]->K) return fail(BM_ERR_ARG, "bm_adj_batch: handles need the same nsteps and checkpoint interval");
		for (int c = 0; c < b; c++) {
			if (hs[c] == hs[b]) return fail(BM_ERR_ARG, "bm_adj_batch: a handle appears twice");
			// the handles run concurrently: each needs its own convolution scratch
			if ((hs[b]->demag && hs[c]->demag == hs[b]->demag) || (hs[b]->ml && hs[c]->ml == hs[b]->ml))
				return fail(BM_ERR_ARG, "bm_adj_batch: handles share a demag or multilayer handle");
		}
	}
	return BM_OK;
}

} // namespace

extern "C" int bm_adj_forward(bm_adj_t o, const float* m0, const float* mt, float* M_final, double* loss, void* stream)
{
	if (!o || !m0 || !mt) return fail(BM_ERR_ARG, "bm_adj_forward: null argument");
	hipStream_t st = (hipStream_t)stream;
	int rc = fwd_begin(o, m0, mt, st);
	for (int s = 0; s < o->nsteps && !rc; s++) rc = fwd_advance(o, s, st);
	if (!rc) rc = fwd_end(o, M_final, st);
	if (rc) return rc;
	BM_HIP(hipStreamSynchronize(st));
	fwd_finish(o, loss);
	BM_HIP(hipGetLastError());
	return BM_OK;
}

extern "C" int bm_adj_backward(bm_adj_t o, float* gMs_out, double* gamp, void* stream)
{
	if (!o || !gMs_out) return fail(BM_ERR_ARG, "bm_adj_backward: null argument");
	if (!o->forward_done) return fail(BM_ERR_STATE, "bm_adj_backward: run bm_adj_forward first");
	hipStream_t st = (hipStream_t)stream;
	int rc = bwd_begin(o, st);
	for (int sg = o->nck - 2; sg >= 0 && !rc; sg--) {
		const int s0 = sg * o->K, s1 = std::min(o->nsteps, s0 + o->K);
		for (int s = s0; s < s1 && !rc; s++) rc = bwd_recompute(o, sg, s, st);
		for (int s = s1 - 1; s >= s0 && !rc; s--) rc = adj_step(o, s, s - s0, st);
	}
	if (!rc) rc = bwd_end(o, gMs_out, st);
	if (rc) return rc;
	BM_HIP(hipStreamSynchronize(st));
	bwd_finish(o, gamp);
	BM_HIP(hipGetLastError());
	return BM_OK;
}

extern "C" int bm_adj_batch_forward(int count, const bm_adj_t* hs, const float* const* m0, const float* const* mt,
	float* const* M_final, double* loss, void* stream)
{
	int rc = batch_check(count, hs);
	if (rc) return rc;
	if (!m0 || !mt) return fail(BM_ERR_ARG, "bm_adj_batch_forward: null argument");
	for (int b = 0; b < count; b++) if (!m0[b] || !mt[b]) return fail(BM_ERR_ARG, "bm_adj_batch_forward: null state");
	hipStream_t st = (hipStream_t)stream;
	if ((rc = batch_fork(hs, count, st))) return rc;
	for (int b = 0; b < count && !rc; b++) rc = fwd_begin(hs[b], m0[b], mt[b], hs[b]->own);
	for (int s = 0; s < hs[0]->nsteps && !rc; s++)
		for (int b = 0; b < count && !rc; b++) rc = fwd_advance(hs[b], s, hs[b]->own);
	for (int b = 0; b < count && !rc; b++) rc = fwd_end(hs[b], M_final ? M_final[b] : nullptr, hs[b]->own);
	if (!rc) rc = batch_join(hs, count, st);
	if (rc) return rc;
	for (int b = 0; b < count; b++) fwd_finish(hs[b], loss ? &loss[b] : nullptr);
	BM_HIP(hipGetLastError());
	return BM_OK;
}

extern "C" int bm_adj_batch_backward(int count, const bm_adj_t* hs, float* const* gMs_out, double* gamp, void* stream)
{
	int rc = batch_check(count, hs);
	if (rc) return rc;
	if (!gMs_out) return fail(BM_ERR_ARG, "bm_adj_batch_backward: null argument");
	for (int b = 0; b < count; b++) {
		if (!gMs_out[b]) return fail(BM_ERR_ARG, "bm_adj_batch_backward: null gradient array");
		if (!hs[b]->forward_done) return fail(BM_ERR_STATE, "bm_adj_batch_backward: run the forward pass first");
	}
	hipStream_t st = (hipStream_t)stream;
	if ((rc = batch_fork(hs, count, st))) return rc;
	for (int b = 0; b < count && !rc; b++) rc = bwd_begin(hs[b], hs[b]->own);
	const int K = hs[0]->K, nsteps = hs[0]->nsteps;
	for (int sg = hs[0]->nck - 2; sg >= 0 && !rc; sg--) {
		const int s0 = sg * K, s1 = std::min(nsteps, s0 + K);
		for (int s = s0; s < s1 && !rc; s++)
			for (int b = 0; b < count && !rc; b++) rc = bwd_recompute(hs[b], sg, s, hs[b]->own);
		for (int s = s1 - 1; s >= s0 && !rc; s--)
			for (int b = 0; b < count && !rc; b++) rc = adj_step(hs[b], s, s - s0, hs[b]->own);
	}
	for (int b = 0; b < count && !rc; b++) rc = bwd_end(hs[b], gMs_out[b], hs[b]->own);
	if (!rc) rc = batch_join(hs, count, st);
	if (rc) return rc;
	for (int b = 0; b < count; b++) bwd_finish(hs[b], gamp ? &gamp[b] : nullptr);
	BM_HIP(hipGetLastError());
	return BM_OK;
}

extern "C" int bm_adj_zeeman_grads(bm_adj_t o, double* out4)
{
	if (!o || !out4) return fail(BM_ERR_ARG, "bm_adj_zeeman_grads: null argument");
	for (int c = 0; c < 4; c++) out4[c] = o->gz[c];
	return BM_OK;
}

// Layer stack for the adjoint (the same contract as bm_ode_set_layers; per-layer Ms comes in
// through the Ms map): the multilayered demag replaces the single-mesh one and BM_MOD_DEMAG
// is switched on; per-layer A, alpha, K1 (NULL keeps the mesh value).
extern "C" int bm_adj_set_layers(bm_adj_t o, bm_mlayer_t ml, const double* A, const double* alpha, const double* K1)
{
	if (!o || !ml) return fail(BM_ERR_ARG, "bm_adj_set_layers: null argument");
	int L, nx, ny;
	mlayer_dims(ml, &L, &nx, &ny, nullptr);
	if (mlayer_mapped(ml)) return fail(BM_ERR_SIZE, "bm_adj_set_layers: the layers must share one planned grid");
	int mp[2];
	bm::mlayer_pbc(ml, mp);
	if ((o->pbc[0] || o->pbc[1] || o->pbc[2]) && (o->pbc[2] || mp[0] != o->pbc[0] || mp[1] != o->pbc[1]))
		return fail(BM_ERR_STATE, "bm_adj_set_layers: the mesh's periodic images differ from the multilayer handle's");
	if (L != o->nz || nx != o->nx || ny != o->ny) return fail(BM_ERR_ARG, "bm_adj_set_layers: the stack's layers must be the mesh's nz planes (same nx, ny)");
	if (L > kMaxAdjLayers) return fail(BM_ERR_SIZE, "bm_adj_set_layers: at most 8 layers");
	const double* src[3] = { A, alpha, K1 };
	for (int q = 0; q < 3; q++) {
		o->lay_set[q] = src[q] != nullptr;
		for (int l = 0; l < kMaxAdjLayers; l++) o->lay[q][l] = src[q] && l < L ? (float)src[q][l] : 0.f;
	}
	o->ml = ml;
	o->p.modules |= BM_MOD_DEMAG;
	o->forward_done = false;
	return BM_OK;
}

// Periodic boundary conditions for the rollout (the contract of bm_ode_set_pbc): exchange
// wraps around along every axis with images, and with demag on the handle given at creation
// must carry the same images.  The adjoint pass needs nothing else: the wrapped Laplacian is
// symmetric and the periodic demag operator is its own adjoint.
extern "C" int bm_adj_set_pbc(bm_adj_t o, int pbc_x, int pbc_y, int pbc_z)
{
	if (!o || pbc_x < 0 || pbc_y < 0 || pbc_z < 0) return fail(BM_ERR_ARG, "bm_adj_set_pbc: bad arguments");
	if (o->ml) {
		// a layer stack: the multilayer handle carries the x, y images (bm_mlayer_create_pbc);
		// its transposed operator stays exact (kernels even or odd along each periodic axis)
		int mp[2];
		bm::mlayer_pbc(o->ml, mp);
		if (pbc_z || mp[0] != pbc_x || mp[1] != pbc_y)
			return fail(BM_ERR_STATE, "bm_adj_set_pbc: a layer stack takes the x, y images of its multilayer handle (bm_mlayer_create_pbc), none along z");
	}
	else if ((o->p.modules & BM_MOD_DEMAG) && (o->demag->pbc[0] != pbc_x || o->demag->pbc[1] != pbc_y || o->demag->pbc[2] != pbc_z))
		return fail(BM_ERR_STATE, "bm_adj_set_pbc: the demag handle must have the same periodic images (bm_demag_create_pbc)");
	o->pbc[0] = pbc_x; o->pbc[1] = pbc_y; o->pbc[2] = pbc_z;
	o->forward_done = false;
	return BM_OK;
}
