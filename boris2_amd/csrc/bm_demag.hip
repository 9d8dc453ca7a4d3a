// Demag convolution for one mesh on MI355X (gfx950): H = IFFT(K . FFT(M)) / (Nx Ny Nz).
//
// Reference behaviour: ConvolutionCUDA::Convolute_2D/3D (Boris/ConvolutionCUDA.h:383/478)
// with the cuFFT pipeline of Boris/ConvolutionDataCUDA.cu:924-1837 and the kernel
// multiplications of Boris/DemagKernelCUDA.cu:13-1635.  The design here is different:
// hand-written LDS FFT passes with the zero-padded planes never read or written and the
// kernel multiply fused into the z pass (3D) or the y pass (2D).
//
// Spectral layout (one complex FP32 buffer S, y fastest — the reference's xy-transposed
// order):  S[c][z][kx][ky], c = component, z < nz (only the planes that hold data),
// kx < NC = Nx/2+1, ky < Ny.
//
//   3D:  P1 x-R2C      M (AoS) -> S[.][z][kx][y<ny]           (rows of a z plane, LDS transpose)
//        P2 y-FFT      S lines (c,z,kx): ny in -> Ny out
//        P3 z-FFT * K * z-IFFT  per (kx,ky) column: nz in -> nz out
//        P4 y-IFFT     Ny in -> ny out
//        P5 x-C2R      -> H (planar or AoS, set or add), scaled by 1/(Nx Ny Nz)
//   2D:  P1, P2' (y-FFT * K * y-IFFT per kx line, 3 components), P5.
//
// Kernel setup runs once on the GPU in FP64 (Newell f/g grids, 27-point stencil or
// far-field series, mirror placement, three real-input FFT passes) and stores FP32
// kernels on the reduced octant, ky fastest.
#include "bm_internal.h"
#include "bm_fft.h"
#include "bm_newell.h"

#include <cstdio>
#include <cstring>
#include <vector>
#include <algorithm>

using namespace bm;

namespace {

constexpr int kBlock = 256;

template <int HX> struct TileY {
	static constexpr int RP = kBlock / plan_T<HX>();
	static constexpr int pw(int v) { int p = 1; while (p * 2 <= v) p *= 2; return p; }
	static constexpr int cap = pw(8400 / (HX + 1)) < 32 ? pw(8400 / (HX + 1)) : 32;
	static constexpr int v = RP > cap ? RP : cap;
};

// ---------------------------------------------------------------------------------------
// P1: x-direction real-to-complex FFT of TY rows of one z plane for one component, via a
// half-length complex FFT (z[k] = x[2k] + i x[2k+1]), then the usual split into the Nc
// half spectrum, transposed into S through LDS.
template <int HX>
__global__ void __launch_bounds__(kBlock) k_fwd_x(const float* __restrict__ M, float2* __restrict__ S,
	int ny, int nz, int Ny, int ntiles)
{
	constexpr int E = Plan<HX>::E, T = plan_T<HX>(), RP = kBlock / T, TY = TileY<HX>::v, NC = HX + 1;
	__shared__ C2<float> tile[TY * NC];
	const int t = threadIdx.x % T, rl = threadIdx.x / T;
	Twiddles<HX, float> tw; tw.init(t);
	C2<float> wk[E];   // exp(-i pi k / HX), k = t + T e
#pragma unroll
	for (int e = 0; e < E; e++) { double s, c; sincospi(-(double)(t + T * e) / HX, &s, &c); wk[e] = mk((float)c, (float)s); }

	const int nwork = 3 * nz * ntiles;
	for (int w = blockIdx.x; w < nwork; w += gridDim.x) {
		const int c = w % 3, z = (w / 3) % nz, y0 = (w / (3 * nz)) * TY;
#pragma unroll 1
		for (int g = 0; g < TY; g += RP) {
			const int yl = g + rl, y = y0 + yl;
			const bool ok = y < ny;
			const float* row = M + ((size_t)z * ny + (ok ? y : 0)) * (size_t)HX * 3 + c;
			C2<float> v[E];
#pragma unroll
			for (int e = 0; e < E; e++) {
				const int k = t + T * e;
				v[e] = (e < E / 2 && ok) ? mk(row[(size_t)(2 * k) * 3], row[(size_t)(2 * k + 1) * 3]) : mk(0.f, 0.f);
			}
			C2<float>* lrow = tile + yl * NC;
			fft<HX, -1, float>(v, lrow, 1, t, tw);
			__syncthreads();
#pragma unroll
			for (int e = 0; e < E; e++) lrow[t + T * e] = v[e];
			__syncthreads();
			C2<float> X[E];
#pragma unroll
			for (int e = 0; e < E; e++) {
				const int k = t + T * e;
				C2<float> zk = v[e], zm = conj(lrow[(HX - k) % HX]);
				C2<float> a = zk + zm, b = zk - zm;
				C2<float> wb = cmul(wk[e], b);              // W^k (Z[k] - conj Z[HX-k])
				X[e] = mk(0.5f * (a.x + wb.y), 0.5f * (a.y - wb.x));   // a/2 - i wb/2
			}
			__syncthreads();
#pragma unroll
			for (int e = 0; e < E; e++) lrow[t + T * e] = X[e];
			if (t == 0) lrow[HX] = mk(v[0].x - v[0].y, 0.f);
		}
		__syncthreads();
		for (int idx = threadIdx.x; idx < TY * NC; idx += kBlock) {
			const int yl = idx % TY, kx = idx / TY, y = y0 + yl;
			if (y < ny) {
				C2<float> a = tile[yl * NC + kx];
				S[(((size_t)c * nz + z) * NC + kx) * (size_t)Ny + y] = make_float2(a.x, a.y);
			}
		}
		__syncthreads();
	}
}

// ---------------------------------------------------------------------------------------
// P5: x-direction complex-to-real inverse FFT.  Output modes: 0 planar set, 1 AoS set,
// 2 AoS add.
template <int HX, int MODE, bool EN>
__global__ void __launch_bounds__(kBlock) k_inv_x(const float2* __restrict__ S, float* __restrict__ out,
	int ny, int nz, int Ny, int ntiles, float scale, const float* __restrict__ M, double* __restrict__ epart)
{
	__shared__ double ered[EN ? kBlock : 1];
	constexpr int E = Plan<HX>::E, T = plan_T<HX>(), RP = kBlock / T, TY = TileY<HX>::v, NC = HX + 1;
	__shared__ C2<float> tile[TY * NC];
	const int t = threadIdx.x % T, rl = threadIdx.x / T;
	Twiddles<HX, float> tw; tw.init(t);
	C2<float> wk[E];   // exp(-i pi k / HX)
#pragma unroll
	for (int e = 0; e < E; e++) { double s, c; sincospi(-(double)(t + T * e) / HX, &s, &c); wk[e] = mk((float)c, (float)s); }
	const int nx = HX;
	const size_t ncell = (size_t)nx * ny * nz;

	const int nwork = 3 * nz * ntiles;
	for (int w = blockIdx.x; w < nwork; w += gridDim.x) {
		const int c = w % 3, z = (w / 3) % nz, y0 = (w / (3 * nz)) * TY;
		double esum = 0.0;
		for (int idx = threadIdx.x; idx < TY * NC; idx += kBlock) {
			const int yl = idx % TY, kx = idx / TY, y = y0 + yl;
			float2 a = (y < ny) ? S[(((size_t)c * nz + z) * NC + kx) * (size_t)Ny + y] : make_float2(0.f, 0.f);
			tile[yl * NC + kx] = mk(a.x, a.y);
		}
		__syncthreads();
#pragma unroll 1
		for (int g = 0; g < TY; g += RP) {
			const int yl = g + rl, y = y0 + yl;
			const bool ok = y < ny;
			C2<float>* lrow = tile + yl * NC;
			C2<float> v[E];
#pragma unroll
			for (int e = 0; e < E; e++) {
				const int k = t + T * e;
				C2<float> xk = lrow[k], xm = conj(lrow[HX - k]);
				C2<float> a = xk + xm, b = xk - xm;
				C2<float> wb = cmulc(b, wk[e]);            // W^-k (X[k] - conj X[HX-k])
				v[e] = mk(a.x - wb.y, a.y + wb.x);          // a + i wb
			}
			fft<HX, 1, float>(v, lrow, 1, t, tw);
			if (ok) {
				const size_t rowc = ((size_t)z * ny + y) * nx;
#pragma unroll
				for (int e = 0; e < E / 2; e++) {
					const int n2 = 2 * (t + T * e);
					float a = v[e].x * scale, b = v[e].y * scale;
					if (MODE == 0) {
						*reinterpret_cast<float2*>(out + (size_t)c * ncell + rowc + n2) = make_float2(a, b);
					}
					else if (MODE == 1) {
						out[(rowc + n2) * 3 + c] = a; out[(rowc + n2 + 1) * 3 + c] = b;
					}
					else {
						out[(rowc + n2) * 3 + c] += a; out[(rowc + n2 + 1) * 3 + c] += b;
					}
					if (EN) esum += (double)(M[(rowc + n2) * 3 + c] * a) + (double)(M[(rowc + n2 + 1) * 3 + c] * b);
				}
			}
		}
		if (EN) {
			// deterministic per-work-item partial of sum M.H
			ered[threadIdx.x] = esum;
			__syncthreads();
			for (int o = kBlock / 2; o > 0; o >>= 1) { if (threadIdx.x < o) ered[threadIdx.x] += ered[threadIdx.x + o]; __syncthreads(); }
			if (threadIdx.x == 0) epart[w] = ered[0];
		}
		__syncthreads();
	}
}

// ---------------------------------------------------------------------------------------
// P2 / P4 (3D): y-direction FFT of lines S[c][z][kx][0..Ny).
template <int NY, int DIR>
__global__ void __launch_bounds__(kBlock) k_y(float2* __restrict__ S, int nlines)
{
	constexpr int E = Plan<NY>::E, T = plan_T<NY>(), LP = kBlock / T;
	__shared__ C2<float> lds[LP * NY];
	const int t = threadIdx.x % T, l = threadIdx.x / T;
	Twiddles<NY, float> tw; tw.init(t);
	const int nwork = (nlines + LP - 1) / LP;
	for (int w = blockIdx.x; w < nwork; w += gridDim.x) {
		const int line = w * LP + l;
		const bool ok = line < nlines;
		float2* p = S + (size_t)(ok ? line : 0) * NY;
		C2<float> v[E];
#pragma unroll
		for (int e = 0; e < E; e++) {
			float2 a = make_float2(0.f, 0.f);
			if (ok && (DIR > 0 || e < E / 2)) a = p[t + T * e];
			v[e] = mk(a.x, a.y);
		}
		fft<NY, DIR, float>(v, lds + l * NY, 1, t, tw);
		if (ok) {
#pragma unroll
			for (int e = 0; e < E; e++)
				if (DIR < 0 || e < E / 2) p[t + T * e] = make_float2(v[e].x, v[e].y);
		}
	}
}

// ---------------------------------------------------------------------------------------
// P3 (3D): per (kx, ky) column, z-FFT of the nz data planes (zero padded to NZ), 6-term
// symmetric kernel multiply, z-IFFT, first nz planes written back.  Kernel symmetries
// (Boris/DemagKernel.cpp:85-168): diagonal even in y and z; Kxy odd in y; Kxz odd in z;
// Kyz odd in y and z.
template <int NZ>
__global__ void __launch_bounds__(kBlock) k_z(float2* __restrict__ S, const float* __restrict__ K,
	int nz, int NC, int Ny, int NYH)
{
	constexpr int E = Plan<NZ>::E, TZ = plan_T<NZ>(), CY = kBlock / TZ, NZH = NZ / 2 + 1;
	__shared__ C2<float> lds[3 * NZ * CY];
	const int col = threadIdx.x % CY, t = threadIdx.x / CY;
	Twiddles<NZ, float> tw; tw.init(t);
	const int nyb = (Ny + CY - 1) / CY;
	const size_t plane = (size_t)NC * Ny, comp = (size_t)nz * plane;
	const size_t kq = (size_t)NZH * NC * NYH;
	for (int w = blockIdx.x; w < NC * nyb; w += gridDim.x) {
		const int kx = w / nyb, ky = (w % nyb) * CY + col;
		const bool ok = ky < Ny;
		float2* p = S + (size_t)kx * Ny + (ok ? ky : 0);
		C2<float> v[3][E];
#pragma unroll
		for (int c = 0; c < 3; c++)
#pragma unroll
			for (int e = 0; e < E; e++) {
				const int z = t + TZ * e;
				float2 a = make_float2(0.f, 0.f);
				if (e < E / 2 && ok && z < nz) a = p[c * comp + z * plane];
				v[c][e] = mk(a.x, a.y);
			}
		fft3<NZ, -1, float>(v[0], v[1], v[2], lds + col, lds + NZ * CY + col, lds + 2 * NZ * CY + col, CY, t, tw);
		const int kyr = ok ? (ky <= Ny / 2 ? ky : Ny - ky) : 0;
		const float sy = (ky <= Ny / 2) ? 1.f : -1.f;
#pragma unroll
		for (int e = 0; e < E; e++) {
			const int kz = t + TZ * e;
			const int kzr = kz <= NZ / 2 ? kz : NZ - kz;
			const float sz = kz <= NZ / 2 ? 1.f : -1.f;
			const size_t ki = ((size_t)kzr * NC + kx) * NYH + kyr;
			const float kxx = K[ki], kyy = K[kq + ki], kzz = K[2 * kq + ki];
			const float kxy = sy * K[3 * kq + ki], kxz = sz * K[4 * kq + ki], kyz = sy * sz * K[5 * kq + ki];
			C2<float> mx = v[0][e], my = v[1][e], mz = v[2][e];
			v[0][e] = mk(kxx * mx.x + kxy * my.x + kxz * mz.x, kxx * mx.y + kxy * my.y + kxz * mz.y);
			v[1][e] = mk(kxy * mx.x + kyy * my.x + kyz * mz.x, kxy * mx.y + kyy * my.y + kyz * mz.y);
			v[2][e] = mk(kxz * mx.x + kyz * my.x + kzz * mz.x, kxz * mx.y + kyz * my.y + kzz * mz.y);
		}
		fft3<NZ, 1, float>(v[0], v[1], v[2], lds + col, lds + NZ * CY + col, lds + 2 * NZ * CY + col, CY, t, tw);
		if (ok) {
#pragma unroll
			for (int c = 0; c < 3; c++)
#pragma unroll
				for (int e = 0; e < E / 2; e++) {
					const int z = t + TZ * e;
					if (z < nz) p[c * comp + z * plane] = make_float2(v[c][e].x, v[c][e].y);
				}
		}
	}
}

// ---------------------------------------------------------------------------------------
// P2' (2D): per kx line, y-FFT of 3 components, 2D kernel multiply (Kxy odd in y), y-IFFT.
template <int NY, int BS>
__global__ void __launch_bounds__(BS) k_y2d(float2* __restrict__ S, const float* __restrict__ K, int NC, int NYH)
{
	constexpr int E = Plan<NY>::E, T = plan_T<NY>(), LP = BS / T;
	__shared__ C2<float> lds[3 * LP * NY];
	const int t = threadIdx.x % T, l = threadIdx.x / T;
	Twiddles<NY, float> tw; tw.init(t);
	const size_t comp = (size_t)NC * NY, kq = (size_t)NC * NYH;
	const int nwork = (NC + LP - 1) / LP;
	for (int w = blockIdx.x; w < nwork; w += gridDim.x) {
		const int kx = w * LP + l;
		const bool ok = kx < NC;
		float2* p = S + (size_t)(ok ? kx : 0) * NY;
		C2<float> v[3][E];
#pragma unroll
		for (int c = 0; c < 3; c++)
#pragma unroll
			for (int e = 0; e < E; e++) {
				float2 a = (ok && e < E / 2) ? p[c * comp + t + T * e] : make_float2(0.f, 0.f);
				v[c][e] = mk(a.x, a.y);
			}
		C2<float>* b0 = lds + l * NY;
		fft3<NY, -1, float>(v[0], v[1], v[2], b0, b0 + LP * NY, b0 + 2 * LP * NY, 1, t, tw);
		const int kxs = ok ? kx : 0;
#pragma unroll
		for (int e = 0; e < E; e++) {
			const int ky = t + T * e;
			const int kyr = ky <= NY / 2 ? ky : NY - ky;
			const float sy = ky <= NY / 2 ? 1.f : -1.f;
			const size_t ki = (size_t)kxs * NYH + kyr;
			const float kxx = K[ki], kyy = K[kq + ki], kzz = K[2 * kq + ki], kxy = sy * K[3 * kq + ki];
			C2<float> mx = v[0][e], my = v[1][e], mz = v[2][e];
			v[0][e] = mk(kxx * mx.x + kxy * my.x, kxx * mx.y + kxy * my.y);
			v[1][e] = mk(kxy * mx.x + kyy * my.x, kxy * mx.y + kyy * my.y);
			v[2][e] = mk(kzz * mz.x, kzz * mz.y);
		}
		fft3<NY, 1, float>(v[0], v[1], v[2], b0, b0 + LP * NY, b0 + 2 * LP * NY, 1, t, tw);
		if (ok) {
#pragma unroll
			for (int c = 0; c < 3; c++)
#pragma unroll
				for (int e = 0; e < E / 2; e++) p[c * comp + t + T * e] = make_float2(v[c][e].x, v[c][e].y);
		}
	}
}

// ---------------------------------------------------------------------------------------
// Energy density  -mu0/(2n) sum M.H from the per-work-item partials of P5.
__global__ void k_energy_final(const double* part, int np, size_t n, float* energy)
{
	__shared__ double red[kBlock];
	double s = 0.0;
	for (int i = threadIdx.x; i < np; i += kBlock) s += part[i];
	red[threadIdx.x] = s;
	__syncthreads();
	for (int o = kBlock / 2; o > 0; o >>= 1) { if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o]; __syncthreads(); }
	if (threadIdx.x == 0) *energy = (float)(-1.256637061435916e-6 * red[0] / (2.0 * (double)n));
}

// =======================================================================================
// FP64 kernel setup
// =======================================================================================

// f (which = 0) or g (which = 1) on the grid of offsets -1..d per axis.
__global__ void k_fill_grid(double* g, int which, int dx, int dy, int dz, double hx, double hy, double hz)
{
	const int gx = dx + 2, gy = dy + 2, gz = dz + 2;
	const size_t total = (size_t)gx * gy * gz;
	for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < total; q += (size_t)gridDim.x * blockDim.x) {
		int i = (int)(q % gx) - 1, j = (int)((q / gx) % gy) - 1, k = (int)(q / ((size_t)gx * gy)) - 1;
		g[q] = which ? newell_g(i * hx, j * hy, k * hz) : newell_f(i * hx, j * hy, k * hz);
	}
}

struct TensorArgs {
	int nx, ny, nz, Nx, Ny, Nz;
	double hx, hy, hz;
	// grids: f_xx, f_yy, f_zz, g_xy, g_xz, g_yz with their dims
	const double* grid[6];
	int gdim[6][2];
	AsymDiag ad[3];
	AsymOff ao[3];
};

// Writes the 6 (3D) or 4 (2D) tensor planes D[q][Nz][Ny][Nx] (-N convention).  The
// mirror images of each first-octant value go to up to 8 places; where an index is zero
// the mirrored writes coincide and the reference's write order decides the sign kept
// (Boris/DemagTFunc.cpp:85-92, 160-167); the loop below keeps the last such write.
__global__ void k_tensor(TensorArgs a, double* D, int is2d)
{
	const size_t NN = (size_t)a.Nx * a.Ny * a.Nz;
	const size_t total = (size_t)a.nx * a.ny * a.nz;
	const int d = kAsymDist;
	// write order of the reference: 3D (Boris/DemagTFunc.cpp:85-92), 2D (:227-230)
	const int mir3[8][3] = { {1,1,1},{-1,1,1},{1,-1,1},{1,1,-1},{-1,-1,1},{-1,1,-1},{1,-1,-1},{-1,-1,-1} };
	const int mir2[4][3] = { {1,1,1},{-1,1,1},{1,-1,1},{-1,-1,1} };
	const int (*mir)[3] = is2d ? mir2 : mir3;
	for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < total; q += (size_t)gridDim.x * blockDim.x) {
		const int i = (int)(q % a.nx), j = (int)((q / a.nx) % a.ny), k = (int)(q / ((size_t)a.nx * a.ny));
		double vd[3], vo[3];
		const bool asym = is2d ? (i >= d || j >= d || i * i + j * j >= d * d)
			: (i >= d || j >= d || k >= d || i * i + j * j + k * k >= d * d);
		const double hx = a.hx, hy = a.hy, hz = a.hz;
		if (asym) {
			vd[0] = -asym_diag(a.ad[0], i * hx, j * hy, k * hz);
			vd[1] = -asym_diag(a.ad[1], j * hy, i * hx, k * hz);
			vd[2] = -asym_diag(a.ad[2], k * hz, j * hy, i * hx);
			vo[0] = -asym_off(a.ao[0], i * hx, j * hy, k * hz);
			vo[1] = is2d ? 0.0 : -asym_off(a.ao[1], i * hx, k * hz, j * hy);
			vo[2] = is2d ? 0.0 : -asym_off(a.ao[2], j * hy, k * hz, i * hx);
		}
		else {
			vd[0] = -stencil27(a.grid[0], a.gdim[0][0], a.gdim[0][1], i, j, k, hx, hy, hz);
			vd[1] = -stencil27(a.grid[1], a.gdim[1][0], a.gdim[1][1], j, i, k, hy, hx, hz);
			vd[2] = -stencil27(a.grid[2], a.gdim[2][0], a.gdim[2][1], k, j, i, hz, hy, hx);
			vo[0] = -stencil27(a.grid[3], a.gdim[3][0], a.gdim[3][1], i, j, k, hx, hy, hz);
			vo[1] = is2d ? 0.0 : -stencil27(a.grid[4], a.gdim[4][0], a.gdim[4][1], i, k, j, hx, hz, hy);
			vo[2] = is2d ? 0.0 : -stencil27(a.grid[5], a.gdim[5][0], a.gdim[5][1], j, k, i, hy, hz, hx);
		}
		const int nm = is2d ? 4 : 8;
		for (int m = 0; m < nm; m++) {
			bool later = false;
			for (int p = m + 1; p < nm; p++)
				if ((mir[p][0] == mir[m][0] || i == 0) && (mir[p][1] == mir[m][1] || j == 0) && (mir[p][2] == mir[m][2] || k == 0)) later = true;
			if (later) continue;
			const int si = mir[m][0], sj = mir[m][1], sk = mir[m][2];
			const size_t idx = (size_t)((si * i + a.Nx) % a.Nx) + (size_t)((sj * j + a.Ny) % a.Ny) * a.Nx
				+ (size_t)((sk * k + a.Nz) % a.Nz) * a.Nx * a.Ny;
			D[idx] = vd[0]; D[NN + idx] = vd[1]; D[2 * NN + idx] = vd[2];
			D[3 * NN + idx] = vo[0] * (si * sj);
			if (!is2d) { D[4 * NN + idx] = vo[1] * (si * sk); D[5 * NN + idx] = vo[2] * (sj * sk); }
		}
	}
}

// Real-input FFT of strided lines in FP64; keeps `keep` outputs, taking the real (part 0)
// or imaginary (part 1) part times `sign`.  Line l = a + na*b reads src[a*sa + b*sb + i*se].
template <int N, typename TO>
__global__ void __launch_bounds__(kBlock) k_line_fft(const double* __restrict__ src, TO* __restrict__ dst, int nlines, int na,
	long long sa, long long sb, long long se, long long da, long long db, long long de, int keep, int part, double sign)
{
	constexpr int E = Plan<N>::E, T = plan_T<N>(), LP = kBlock / T;
	__shared__ C2<double> lds[LP * N];
	const int t = threadIdx.x % T, l = threadIdx.x / T;
	Twiddles<N, double> tw; tw.init(t);
	const int nwork = (nlines + LP - 1) / LP;
	for (int w = blockIdx.x; w < nwork; w += gridDim.x) {
		const int line = w * LP + l;
		const bool ok = line < nlines;
		const int a = ok ? line % na : 0, b = ok ? line / na : 0;
		const double* p = src + a * sa + b * sb;
		C2<double> v[E];
#pragma unroll
		for (int e = 0; e < E; e++) v[e] = mk(ok ? p[(t + T * e) * se] : 0.0, 0.0);
		fft<N, -1, double>(v, lds + l * N, 1, t, tw);
		if (ok) {
			TO* q = dst + a * da + b * db;
#pragma unroll
			for (int e = 0; e < E; e++) {
				const int i = t + T * e;
				if (i < keep) q[i * de] = (TO)(sign * (part ? v[e].y : v[e].x));
			}
		}
	}
}

// ---------------------------------------------------------------------------------------
// Size dispatch.
#define BM_FOR_X_SIZES(M) M(2) M(4) M(6) M(8) M(12) M(16) M(24) M(32) M(48) M(64) M(96) M(128) M(192) M(256) M(384) M(512) M(768) M(1024) M(2048)
#define BM_FOR_Z_SIZES(M) M(4) M(8) M(16) M(32) M(64) M(128)

bool x_supported(int n) { switch (n) {
#define C_(V_) case V_:
	BM_FOR_X_SIZES(C_)
#undef C_
	return true; default: return false; } }
bool line_supported(int n) { return x_supported(n) || n == 1536 || n == 4096; }
bool z_supported(int n) { switch (n) {
#define C_(V_) case V_:
	BM_FOR_Z_SIZES(C_)
#undef C_
	return true; default: return false; } }

int grid_for(int nwork, int ncu, int per_cu) { return std::max(1, std::min(nwork, ncu * per_cu)); }

int launch_fwd_x(int HX, const float* M, float2* S, int ny, int nz, int Ny, int ncu, hipStream_t st)
{
	switch (HX) {
#define C_(V_) case V_: { int TY = TileY<V_>::v, nt = (ny + TY - 1) / TY; \
		hipLaunchKernelGGL(k_fwd_x<V_>, dim3(grid_for(3 * nz * nt, ncu, 4)), dim3(kBlock), 0, st, M, S, ny, nz, Ny, nt); return BM_OK; }
	BM_FOR_X_SIZES(C_)
#undef C_
	}
	return fail(BM_ERR_SIZE, "no x plan");
}

int launch_inv_x(int HX, int mode, const float2* S, float* out, int ny, int nz, int Ny, float scale, int ncu, hipStream_t st,
	const float* M = nullptr, double* epart = nullptr)
{
	switch (HX) {
#define C_(V_) case V_: { int TY = TileY<V_>::v, nt = (ny + TY - 1) / TY; dim3 g(grid_for(3 * nz * nt, ncu, 4)); \
		if (mode == 0) hipLaunchKernelGGL((k_inv_x<V_, 0, false>), g, dim3(kBlock), 0, st, S, out, ny, nz, Ny, nt, scale, M, epart); \
		else if (mode == 1 && !epart) hipLaunchKernelGGL((k_inv_x<V_, 1, false>), g, dim3(kBlock), 0, st, S, out, ny, nz, Ny, nt, scale, M, epart); \
		else if (mode == 1) hipLaunchKernelGGL((k_inv_x<V_, 1, true>), g, dim3(kBlock), 0, st, S, out, ny, nz, Ny, nt, scale, M, epart); \
		else if (!epart) hipLaunchKernelGGL((k_inv_x<V_, 2, false>), g, dim3(kBlock), 0, st, S, out, ny, nz, Ny, nt, scale, M, epart); \
		else hipLaunchKernelGGL((k_inv_x<V_, 2, true>), g, dim3(kBlock), 0, st, S, out, ny, nz, Ny, nt, scale, M, epart); return BM_OK; }
	BM_FOR_X_SIZES(C_)
#undef C_
	}
	return fail(BM_ERR_SIZE, "no x plan");
}

int inv_x_nwork(int HX, int ny, int nz)
{
	switch (HX) {
#define C_(V_) case V_: { int TY = TileY<V_>::v; return 3 * nz * ((ny + TY - 1) / TY); }
	BM_FOR_X_SIZES(C_)
#undef C_
	}
	return 0;
}

int launch_y(int NY, int dir, float2* S, int nlines, int ncu, hipStream_t st)
{
	switch (NY) {
#define C_(V_) case V_: { int LP = kBlock / plan_T<V_>(); dim3 g(grid_for((nlines + LP - 1) / LP, ncu, 8)); \
		if (dir < 0) hipLaunchKernelGGL((k_y<V_, -1>), g, dim3(kBlock), 0, st, S, nlines); \
		else hipLaunchKernelGGL((k_y<V_, 1>), g, dim3(kBlock), 0, st, S, nlines); return BM_OK; }
	BM_FOR_X_SIZES(C_)
#undef C_
	}
	return fail(BM_ERR_SIZE, "no y plan");
}

int launch_z(int NZ, float2* S, const float* K, int nz, int NC, int Ny, int NYH, int ncu, hipStream_t st)
{
	switch (NZ) {
#define C_(V_) case V_: { int CY = kBlock / plan_T<V_>(); int nw = NC * ((Ny + CY - 1) / CY); \
		hipLaunchKernelGGL(k_z<V_>, dim3(grid_for(nw, ncu, 8)), dim3(kBlock), 0, st, S, K, nz, NC, Ny, NYH); return BM_OK; }
	BM_FOR_Z_SIZES(C_)
#undef C_
	}
	return fail(BM_ERR_SIZE, "no z plan");
}

int launch_y2d(int NY, float2* S, const float* K, int NC, int NYH, int ncu, hipStream_t st)
{
	switch (NY) {
#define C_(V_) case V_: { constexpr int BS = Plan<V_>::E > 8 ? 128 : 256; int LP = BS / plan_T<V_>(); \
		hipLaunchKernelGGL((k_y2d<V_, BS>), dim3(grid_for((NC + LP - 1) / LP, ncu, 8)), dim3(BS), 0, st, S, K, NC, NYH); return BM_OK; }
	BM_FOR_X_SIZES(C_)
#undef C_
	}
	return fail(BM_ERR_SIZE, "no y plan");
}

template <typename TO>
int launch_line_fft(int N, const double* src, TO* dst, int nlines, int na, long long sa, long long sb, long long se,
	long long da, long long db, long long de, int keep, int part, double sign, hipStream_t st)
{
	switch (N) {
#define C_(V_) case V_: { int LP = kBlock / plan_T<V_>(); \
		hipLaunchKernelGGL((k_line_fft<V_, TO>), dim3(std::max(1, std::min((nlines + LP - 1) / LP, 8192))), dim3(kBlock), 0, st, \
			src, dst, nlines, na, sa, sb, se, da, db, de, keep, part, sign); return BM_OK; }
	BM_FOR_X_SIZES(C_)
	C_(1536) C_(4096)
#undef C_
	}
	return fail(BM_ERR_SIZE, "no setup FFT plan");
}

// ---------------------------------------------------------------------------------------
// Kernel computation on the GPU (FP64 tensor and transforms, FP32 result).
int compute_kernels(bm_demag_s* d, hipStream_t st)
{
	const int nx = d->nx, ny = d->ny, nz = d->nz, Nx = d->Nx, Ny = d->Ny, Nz = d->Nz;
	const int NC = d->NC, NYH = d->NYH, NZH = d->NZH;
	double hm = std::max(d->h[0], std::max(d->h[1], d->h[2]));
	double hx = d->h[0] / hm, hy = d->h[1] / hm, hz = d->h[2] / hm;
	const int dd = kAsymDist;
	int dx = std::min(nx, dd), dy = std::min(ny, dd), dz = std::min(nz, dd);

	TensorArgs a;
	std::memset(&a, 0, sizeof(a));
	a.nx = nx; a.ny = ny; a.nz = nz; a.Nx = Nx; a.Ny = Ny; a.Nz = Nz; a.hx = hx; a.hy = hy; a.hz = hz;
	// grid extents (offsets -1..d) and the argument scalings of each basis-function grid
	const int gd[6][3] = { {dx,dy,dz},{dy,dx,dz},{dz,dy,dx},{dx,dy,dz},{dx,dz,dy},{dy,dz,dx} };
	const double gh[6][3] = { {hx,hy,hz},{hy,hx,hz},{hz,hy,hx},{hx,hy,hz},{hx,hz,hy},{hy,hz,hx} };
	const int ngrid = d->is2d ? 4 : 6;
	double* grids[6] = { nullptr };
	size_t NN = (size_t)Nx * Ny * Nz;
	const int ncomp = d->is2d ? 4 : 6;
	double* D = nullptr; double* T1 = nullptr; double* T2 = nullptr;
	int rc = BM_OK;
	auto cleanup = [&]() {
		(void)hipStreamSynchronize(st);
		for (int q = 0; q < 6; q++) if (grids[q]) (void)hipFree(grids[q]);
		if (D) (void)hipFree(D); if (T1) (void)hipFree(T1); if (T2) (void)hipFree(T2);
	};
	for (int q = 0; q < ngrid; q++) {
		size_t cnt = (size_t)(gd[q][0] + 2) * (gd[q][1] + 2) * (gd[q][2] + 2);
		if (hipMalloc(&grids[q], cnt * sizeof(double)) != hipSuccess) { cleanup(); return fail(BM_ERR_NOMEM, "tensor grids"); }
		hipLaunchKernelGGL(k_fill_grid, dim3((unsigned)std::min<size_t>((cnt + 255) / 256, 4096)), dim3(256), 0, st,
			grids[q], q >= 3 ? 1 : 0, gd[q][0], gd[q][1], gd[q][2], gh[q][0], gh[q][1], gh[q][2]);
		a.grid[q] = grids[q]; a.gdim[q][0] = gd[q][0] + 2; a.gdim[q][1] = gd[q][1] + 2;
	}
	asym_diag_init(a.ad[0], hx, hy, hz); asym_diag_init(a.ad[1], hy, hx, hz); asym_diag_init(a.ad[2], hz, hy, hx);
	asym_off_init(a.ao[0], hx, hy, hz); asym_off_init(a.ao[1], hx, hz, hy); asym_off_init(a.ao[2], hy, hz, hx);

	if (hipMalloc(&D, NN * ncomp * sizeof(double)) != hipSuccess) { cleanup(); return fail(BM_ERR_NOMEM, "tensor"); }
	if (hipMemsetAsync(D, 0, NN * ncomp * sizeof(double), st) != hipSuccess) { cleanup(); return fail(BM_ERR_HIP, "memset"); }
	{
		size_t total = (size_t)nx * ny * nz;
		hipLaunchKernelGGL(k_tensor, dim3((unsigned)std::min<size_t>((total + 127) / 128, 65535)), dim3(128), 0, st, a, D, d->is2d ? 1 : 0);
	}
	if (hipMalloc(&T1, (size_t)NC * Ny * Nz * sizeof(double)) != hipSuccess) { cleanup(); return fail(BM_ERR_NOMEM, "tensor scratch"); }
	if (!d->is2d && hipMalloc(&T2, (size_t)NC * NYH * Nz * sizeof(double)) != hipSuccess) { cleanup(); return fail(BM_ERR_NOMEM, "tensor scratch"); }

	// per-component part selection along x, y, z and final sign (Boris/DemagKernel.cpp:388-468
	// for 3D; :272-307 for 2D): 0 = real part, 1 = imaginary part
	for (int q = 0; q < ncomp && rc == BM_OK; q++) {
		const double* Dq = D + (size_t)q * NN;
		float* Kq = d->K + (size_t)q * (d->is2d ? (size_t)NC * NYH : (size_t)NZH * NC * NYH);
		if (d->is2d) {
			int px = (q == 3) ? 1 : 0, py = (q == 3) ? 1 : 0;
			double sgn = (q == 3) ? -1.0 : 1.0;
			rc = launch_line_fft<double>(Nx, Dq, T1, Ny, Ny, Nx, 0, 1, NC, 0, 1, NC, px, 1.0, st);
			if (rc == BM_OK) rc = launch_line_fft<float>(Ny, T1, Kq, NC, NC, 1, 0, NC, NYH, 0, 1, NYH, py, sgn, st);
		}
		else {
			static const int part[6][3] = { {0,0,0},{0,0,0},{0,0,0},{1,1,0},{1,0,1},{0,1,1} };
			double sgn = q >= 3 ? -1.0 : 1.0;
			// x: line (y, z) -> T1[z][y][kx]
			rc = launch_line_fft<double>(Nx, Dq, T1, Ny * Nz, Ny * Nz, Nx, 0, 1, NC, 0, 1, NC, part[q][0], 1.0, st);
			// y: line (kx, z) -> T2[z][ky][kx]
			if (rc == BM_OK) rc = launch_line_fft<double>(Ny, T1, T2, NC * Nz, NC, 1, (long long)NC * Ny, NC, 1, (long long)NC * NYH, NC, NYH, part[q][1], 1.0, st);
			// z: line (kx, ky) -> K[kz][kx][ky]
			if (rc == BM_OK) rc = launch_line_fft<float>(Nz, T2, Kq, NC * NYH, NC, 1, NC, (long long)NC * NYH, NYH, 1, (long long)NC * NYH, NZH, part[q][2], sgn, st);
		}
	}
	if (rc == BM_OK) { hipError_t e = hipStreamSynchronize(st); if (e != hipSuccess) rc = hip_fail(e, "kernel setup"); }
	cleanup();
	return rc;
}

} // namespace

// =======================================================================================
// C-ABI
// =======================================================================================

// The convolution passes up to (not including) the final x C2R pass.
static int forward_and_multiply(bm_demag_s* d, const float* M, hipStream_t st)
{
	const int HX = d->nx, nlines = 3 * d->nz * d->NC;
	prof_mark_begin(PK_FWD_X, st);
	int rc = launch_fwd_x(HX, M, d->S, d->ny, d->nz, d->Ny, d->ncu, st);
	prof_mark_end(PK_FWD_X, st);
	if (rc) return rc;
	if (d->is2d) {
		prof_mark_begin(PK_Y2D, st);
		rc = launch_y2d(d->Ny, d->S, d->K, d->NC, d->NYH, d->ncu, st);
		prof_mark_end(PK_Y2D, st);
		return rc;
	}
	prof_mark_begin(PK_Y_FWD, st);
	rc = launch_y(d->Ny, -1, d->S, nlines, d->ncu, st);
	prof_mark_end(PK_Y_FWD, st);
	if (rc) return rc;
	prof_mark_begin(PK_Z, st);
	rc = launch_z(d->Nz, d->S, d->K, d->nz, d->NC, d->Ny, d->NYH, d->ncu, st);
	prof_mark_end(PK_Z, st);
	if (rc) return rc;
	prof_mark_begin(PK_Y_INV, st);
	rc = launch_y(d->Ny, 1, d->S, nlines, d->ncu, st);
	prof_mark_end(PK_Y_INV, st);
	return rc;
}

int bm::demag_planar(bm_demag_s* d, const float* M, float* H3n, hipStream_t st)
{
	int rc = forward_and_multiply(d, M, st);
	if (rc) return rc;
	const float scale = (float)(1.0 / ((double)d->Nx * d->Ny * d->Nz));
	prof_mark_begin(PK_INV_X, st);
	rc = launch_inv_x(d->nx, 0, d->S, H3n, d->ny, d->nz, d->Ny, scale, d->ncu, st);
	prof_mark_end(PK_INV_X, st);
	return rc;
}

extern "C" int bm_demag_create(int nx, int ny, int nz, double hx, double hy, double hz, int flags, bm_demag_t* out)
{
	(void)flags;
	if (!out || nx < 2 || ny < 1 || nz < 1 || !(hx > 0) || !(hy > 0) || !(hz > 0)) return fail(BM_ERR_ARG, "bm_demag_create: bad arguments");
	*out = nullptr;
	if (nx % 2) return fail(BM_ERR_SIZE, "bm_demag_create: nx must be even");
	bm_demag_s* d = new bm_demag_s();
	d->nx = nx; d->ny = ny; d->nz = nz; d->h[0] = hx; d->h[1] = hy; d->h[2] = hz;
	d->Nx = 2 * nx; d->Ny = 2 * ny;
	d->Nz = 1;
	if (nz > 1) { while (d->Nz < 2 * nz) d->Nz *= 2; }
	d->is2d = (nz == 1);
	d->NC = d->Nx / 2 + 1; d->NYH = d->Ny / 2 + 1; d->NZH = d->Nz / 2 + 1;
	// nx is the half-length complex FFT of the x pass; Nx, Ny, Nz also need FP64 setup plans
	if (!x_supported(nx) || !x_supported(d->Ny) || !line_supported(d->Nx) || (!d->is2d && !z_supported(d->Nz))) {
		delete d;
		return fail(BM_ERR_SIZE, "bm_demag_create: no FFT plan for this mesh size (supported: nx in {2..2048 of 2^a 3^b plans}, Ny = 2ny likewise, Nz <= 128)");
	}
	d->ncu = device_cu_count();
	size_t sbytes = (size_t)3 * nz * d->NC * d->Ny * sizeof(float2);
	d->kcount = d->is2d ? 4LL * d->NC * d->NYH : 6LL * d->NZH * d->NC * d->NYH;
	if (hipMalloc(&d->S, sbytes) != hipSuccess) { delete d; return fail(BM_ERR_NOMEM, "bm_demag_create: spectra"); }
	if (hipMalloc(&d->K, d->kcount * sizeof(float)) != hipSuccess) { (void)hipFree(d->S); delete d; return fail(BM_ERR_NOMEM, "bm_demag_create: kernels"); }
	d->ecap = std::max(1, inv_x_nwork(nx, ny, nz));
	if (hipMalloc(&d->ebuf, (size_t)d->ecap * sizeof(double)) != hipSuccess) { (void)hipFree(d->S); (void)hipFree(d->K); delete d; return fail(BM_ERR_NOMEM, "bm_demag_create: energy"); }
	d->bytes = (long long)sbytes + d->kcount * 4 + (long long)d->ecap * 8;
	hipStream_t st = nullptr;
	int rc = compute_kernels(d, st);
	if (rc) { bm_demag_destroy(d); return rc; }
	*out = d;
	return BM_OK;
}

extern "C" int bm_demag_destroy(bm_demag_t d)
{
	if (!d) return BM_OK;
	if (d->S) (void)hipFree(d->S);
	if (d->K) (void)hipFree(d->K);
	if (d->ebuf) (void)hipFree(d->ebuf);
	delete d;
	return BM_OK;
}

extern "C" int bm_demag_info(bm_demag_t d, int* N3, long long* bytes)
{
	if (!d) return fail(BM_ERR_ARG, "null handle");
	if (N3) { N3[0] = d->Nx; N3[1] = d->Ny; N3[2] = d->Nz; }
	if (bytes) *bytes = d->bytes;
	return BM_OK;
}

extern "C" int bm_demag_convolute(bm_demag_t d, const float* M, float* Heff, int clear_out, float* energy, void* stream)
{
	if (!d || !M || !Heff) return fail(BM_ERR_ARG, "bm_demag_convolute: null argument");
	hipStream_t st = (hipStream_t)stream;
	const int HX = d->nx;
	int rc = forward_and_multiply(d, M, st);
	if (rc) return rc;
	const float scale = (float)(1.0 / ((double)d->Nx * d->Ny * d->Nz));
	if (energy) {
		const int np = inv_x_nwork(HX, d->ny, d->nz);
		if (np > d->ecap) return fail(BM_ERR_STATE, "energy partial buffer too small");
		prof_mark_begin(PK_INV_X, st);
		rc = launch_inv_x(HX, clear_out ? 1 : 2, d->S, Heff, d->ny, d->nz, d->Ny, scale, d->ncu, st, M, d->ebuf);
		prof_mark_end(PK_INV_X, st);
		if (rc) return rc;
		size_t n = (size_t)d->nx * d->ny * d->nz;
		hipLaunchKernelGGL(k_energy_final, dim3(1), dim3(kBlock), 0, st, d->ebuf, np, n, energy);
	}
	else {
		prof_mark_begin(PK_INV_X, st);
		rc = launch_inv_x(HX, clear_out ? 1 : 2, d->S, Heff, d->ny, d->nz, d->Ny, scale, d->ncu, st);
		prof_mark_end(PK_INV_X, st);
		if (rc) return rc;
	}
	BM_HIP(hipGetLastError());
	return BM_OK;
}

extern "C" int bm_demag_convolute_planar(bm_demag_t d, const float* M, float* H3n, void* stream)
{
	if (!d || !M || !H3n) return fail(BM_ERR_ARG, "bm_demag_convolute_planar: null argument");
	int rc = bm::demag_planar(d, M, H3n, (hipStream_t)stream);
	if (rc) return rc;
	BM_HIP(hipGetLastError());
	return BM_OK;
}

extern "C" int bm_demag_copy_kernels(bm_demag_t d, float* host_out, long long count)
{
	if (!d || !host_out || count < d->kcount) return fail(BM_ERR_ARG, "bm_demag_copy_kernels: bad arguments");
	BM_HIP(hipMemcpy(host_out, d->K, d->kcount * sizeof(float), hipMemcpyDeviceToHost));
	return BM_OK;
}
