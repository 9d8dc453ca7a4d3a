// Demag convolution for one mesh on MI355X (gfx950): H = IFFT(K . FFT(M)) / (Nx Ny Nz).
//
// Reference behaviour: ConvolutionCUDA::Convolute_2D/3D (Boris/ConvolutionCUDA.h:383/478)
// with the cuFFT pipeline of Boris/ConvolutionDataCUDA.cu:924-1837 and the kernel
// multiplications of Boris/DemagKernelCUDA.cu:13-1635.  The design here is different:
// hand-written LDS FFT passes with the zero-padded planes never read or written and the
// kernel multiply fused into the z pass (3D) or the y pass (2D).
//
// Spectral layouts (complex FP32), c = component, z < nz (only the planes that hold data),
// kx < NC = Nx/2+1:
//   Sx  x-side spectra of the ny data rows, row-tiled: Sx[c][z][y/8][kx][y%8] (bm_xpass.h);
//   S   full spectra (3D), y fastest — the reference's xy-transposed order: S[c][z][kx][ky<Ny].
//
//   3D:  P1 x-R2C      M (AoS) -> Sx                      (rows of a z plane through LDS)
//        P2 y-FFT      Sx lines (c,z,kx): ny in -> S lines: Ny out
//        P3 z-FFT * K * z-IFFT  per (kx,ky) column of S: nz in -> nz out
//        P4 y-IFFT     S: Ny in -> Sx: ny out
//        P5 x-C2R      Sx -> H (planar or AoS, set or add), scaled by 1/(Nx Ny Nz)
//   2D:  P1, P2' (y-FFT * K * y-IFFT per kx line of Sx, 3 components, in place), P5.
//
// Kernel setup runs once on the GPU in FP64 (Newell f/g grids, 27-point stencil or
// far-field series, mirror placement, three real-input FFT passes) and stores FP32
// kernels on the reduced octant (3D: kz fastest; 2D: ky fastest).
#include "bm_internal.h"
#include "bm_fft.h"
#include "bm_xpass.h"
#include "bm_newell.h"

#include <cstdio>
#include <cstring>
#include <vector>
#include <algorithm>
#include <cstdlib>
#include <mutex>
#include <string>
#include <type_traits>
#include <string>

using namespace bm;

namespace {

constexpr int kBlock = 256;

// ---------------------------------------------------------------------------------------
// Batched launches (many meshes of one pass size in one launch): grid row blockIdx.y takes
// its mesh's pointers and sizes from a device table (FwdArgs / PassBatch, bm_internal.h);
// the single-mesh launches pass null.

// ---------------------------------------------------------------------------------------
// P1: x-direction real-to-complex FFT of TY rows of one z plane (all three components, so M
// is read from HBM once), through LDS into Sx.  With ADV it also performs the
// RKDP54 advance on the way: sM1 = M, Mout = M + E0 dT/5, and transforms Mout.
// FULL (periodic x, N.x = nx): a row holds 2 HX cells and every packed pair carries data;
// otherwise a row holds HX cells and the upper half of the packed pairs is the zero padding.
// GEN (N.x larger than 2 nx, any nx <= HX, odd included): a row holds a.nx cells, loaded one
// float at a time (rows of odd length are not 8-byte aligned) with the cells past nx zero.

template <int HX, int TY, bool ADV, bool FULL = false, bool GEN = false>
__global__ void __launch_bounds__(kBlock) k_fwd_x(FwdArgs a, const FwdArgs* __restrict__ bt)
{
	static_assert(!(GEN && (ADV || FULL)), "the general row length has no fused advance and no periodic x");
	if (bt) a = bt[blockIdx.y];     // batched: one mesh per grid row
	constexpr int E = Plan<HX>::E, T = plan_T<HX>(), RP = kBlock / T, RPITCH = xrow_pitch<HX>(), NG = TY / RP;
	constexpr int EL = FULL ? E : E / 2, NXR = FULL ? 2 * HX : HX;   // pairs per thread, cells per row
	const int nxr = GEN ? a.nx : NXR;
	__shared__ C2<float> tile[TY * RPITCH];
	XThread<HX> th; th.init();
	const int nwork = a.nz * a.ntiles;
	for (int w = blockIdx.x; w < nwork; w += gridDim.x) {
		const int z = w % a.nz, y0 = (w / a.nz) * TY;
		// cells 2k and 2k+1 of each of this thread's rows: 6 floats = 3 aligned float2
		float2 m[NG][EL][3];
#pragma unroll
		for (int g = 0; g < NG; g++) {
			const int y = y0 + g * RP + th.rl;
			const bool ok = y < a.ny;
			const size_t rowoff = ((size_t)z * a.ny + (ok ? y : 0)) * (size_t)nxr * 3;
			const float2* row = reinterpret_cast<const float2*>(a.M + rowoff);
#pragma unroll
			for (int e = 0; e < EL; e++) {
				const int k = th.t + T * e;
				if constexpr (GEN) {
					const float* c0 = a.M + rowoff + 6 * (size_t)k;
					const bool v0 = ok && 2 * k < nxr, v1 = ok && 2 * k + 1 < nxr;
					float f[6];
#pragma unroll
					for (int q = 0; q < 3; q++) { f[q] = v0 ? c0[q] : 0.f; f[3 + q] = v1 ? c0[3 + q] : 0.f; }
					m[g][e][0] = make_float2(f[0], f[1]); m[g][e][1] = make_float2(f[2], f[3]); m[g][e][2] = make_float2(f[4], f[5]);
					continue;
				}
#pragma unroll
				for (int q = 0; q < 3; q++) m[g][e][q] = ok ? row[3 * k + q] : make_float2(0.f, 0.f);
				if (ADV && ok) {
					const float2* er = reinterpret_cast<const float2*>(a.E0 + rowoff);
					float2* sr = reinterpret_cast<float2*>(a.sM1 + rowoff);
					float2* mr = reinterpret_cast<float2*>(a.Mout + rowoff);
#pragma unroll
					for (int q = 0; q < 3; q++) {
						const float2 e0 = er[3 * k + q];
						sr[3 * k + q] = m[g][e][q];
						m[g][e][q] = make_float2(m[g][e][q].x + e0.x * a.dT5, m[g][e][q].y + e0.y * a.dT5);
						mr[3 * k + q] = m[g][e][q];
					}
				}
			}
		}
#pragma unroll 1
		for (int c = 0; c < 3; c++) {
#pragma unroll
			for (int g = 0; g < NG; g++) {
				C2<float> v[E];
#pragma unroll
				for (int e = 0; e < E; e++) {
					if (e < EL) {
						// component c of cell 2k is float c of the 6; of cell 2k+1 float 3 + c
						const float2 p = m[g][e][0], q = m[g][e][1], r = m[g][e][2];
						const float x0 = (c == 0) ? p.x : (c == 1) ? p.y : q.x;
						const float x1 = (c == 0) ? q.y : (c == 1) ? r.x : r.y;
						v[e] = mk(x0, x1);
					}
					else v[e] = mk(0.f, 0.f);
				}
				r2c_row<HX>(th, tile + (g * RP + th.rl) * RPITCH, v);
			}
			__syncthreads();
			tile_to_S<HX, TY>(tile, a.Sx, c, z + a.zoff, y0, a.ny, a.nzl ? a.nzl : a.nz, a.ntl);
			__syncthreads();
		}
	}
}

// ---------------------------------------------------------------------------------------
// P5: x-direction complex-to-real inverse FFT.  Output modes: 0 planar set, 1 AoS set,
// 2 AoS add; EN also forms per-work-item partial sums of M.H for the energy.  FULL as in P1:
// all 2 HX outputs of a row are cells (periodic x), otherwise the first HX; GEN as in P1: the
// first nxr (<= HX) outputs of a row are its cells, stored one float at a time.
template <int HX, int MODE, bool EN, int TY, bool FULL = false, bool GEN = false>
__global__ void __launch_bounds__(kBlock) k_inv_x(const float2* __restrict__ Sx, float* __restrict__ out,
	int ny, int nz, int ntl, int ntiles, float scale, const float* __restrict__ M, double* __restrict__ epart,
	int nxr, int zoff, int nzl, const PassBatch* __restrict__ bt)
{
	static_assert(!(GEN && FULL), "the general row length has no periodic x");
	if (bt) { const PassBatch& b = bt[blockIdx.y]; Sx = b.Sx; out = b.out; ny = b.ny; nz = b.nz; ntl = b.ntl; ntiles = b.ntiles; scale = b.scale; }
	__shared__ double ered[EN ? kBlock : 1];
	constexpr int E = Plan<HX>::E, T = plan_T<HX>(), RP = kBlock / T, RPITCH = xrow_pitch<HX>();
	__shared__ C2<float> tile[TY * RPITCH];
	XThread<HX> th; th.init();
	constexpr int EL = FULL ? E : E / 2;
	const int nx = GEN ? nxr : FULL ? 2 * HX : HX;
	const size_t ncell = (size_t)nx * ny * nz;
	// the next work item's tile is loaded into registers while this one is transformed
	XTileRegs<HX, TY> pre;
	const int nwork = 3 * nz * ntiles;
	const int nzs = nzl ? nzl : nz;      // planes per component of the Sx layout
	auto fetch = [&](int wf) { pre.load(Sx, wf % 3, zoff + (wf / 3) % nz, (wf / (3 * nz)) * TY, ny, nzs, ntl); };
	if ((int)blockIdx.x < nwork) fetch(blockIdx.x);
	for (int w = blockIdx.x; w < nwork; w += gridDim.x) {
		const int c = w % 3, z = (w / 3) % nz, y0 = (w / (3 * nz)) * TY;
		double esum = 0.0;
		__syncthreads();
		pre.to_tile(tile);
		__syncthreads();
		if (w + (int)gridDim.x < nwork) fetch(w + gridDim.x);
#pragma unroll 1
		for (int g = 0; g < TY; g += RP) {
			const int y = y0 + g + th.rl;
			const bool ok = y < ny;
			C2<float> v[E];
			c2r_row<HX>(th, tile + (g + th.rl) * RPITCH, v);
			if (ok) {
				const size_t rowc = ((size_t)z * ny + y) * nx;
#pragma unroll
				for (int e = 0; e < EL; e++) {
					const int n2 = 2 * (th.t + T * e);
					float a = v[e].x * scale, b = v[e].y * scale;
					if constexpr (GEN) {
						const bool v0 = n2 < nx, v1 = n2 + 1 < nx;
						float* o0 = MODE == 0 ? out + (size_t)c * ncell + rowc + n2 : out + (rowc + n2) * 3 + c;
						float* o1 = MODE == 0 ? o0 + 1 : o0 + 3;
						if (v0) *o0 = MODE == 2 ? *o0 + a : a;
						if (v1) *o1 = MODE == 2 ? *o1 + b : b;
						if (EN) esum += (v0 ? (double)(M[(rowc + n2) * 3 + c] * a) : 0.0) + (v1 ? (double)(M[(rowc + n2 + 1) * 3 + c] * b) : 0.0);
						continue;
					}
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
	}
}

// ---------------------------------------------------------------------------------------
// P2 / P4 (3D): y-direction FFT of lines (c, z, kx), line = (c nz + z) NC + kx.  DIR -1:
// the ny = Ny/2 data rows of a line come from the row-tiled Sx and its Ny transformed values
// go to line `line` of S; DIR +1: S line in, first ny outputs to Sx.  A block takes LP
// consecutive lines, one per group of T lanes: a wave moves its line's S values as whole
// runs and its Sx values as kLT-value pieces, the block's waves together covering runs of
// LP * kLT consecutive Sx values.  (Measured at 512x512x32: staging the Sx side through LDS
// so that each wave moves whole runs, or interleaving the lanes over the lines, is slower.)
// FULL (periodic y, N.y = ny): all NY values of a line are data rows.
// nyd: data rows of a line (ny; N.y when periodic); N.y may exceed 2 ny, the rows between
// ny and N.y / 2 are zeros that are never stored.  V: radix schedule (plan variant) of NY.
template <int NY, int DIR, bool FULL = false, int V = 0>
__global__ void __launch_bounds__(kBlock) k_y(float2* __restrict__ S, float2* __restrict__ Sx, int nlines, int Sp, int NC, int ntl,
	int nyd, const PassBatch* __restrict__ bt)
{
	if (bt) { const PassBatch& b = bt[blockIdx.y]; S = b.S; Sx = b.Sx; nlines = b.nlines; Sp = b.Sp; NC = b.NC; ntl = b.ntl; nyd = b.ny; }
	constexpr int PK = plan_key(NY, V);
	constexpr int E = Plan<PK>::E, T = plan_T<PK>(), LP = kBlock / T, LL = lds_len<NY>(), EL = FULL ? E : E / 2;
	__shared__ C2<float> lds[LP * LL];
	const int t = threadIdx.x % T, l = threadIdx.x / T;
	Twiddles<PK, float> tw; tw.init(t);
	const int nwork = (nlines + LP - 1) / LP;
	// the next line group is loaded into registers while this one is transformed
	float2 pre[E];
	auto fetch = [&](int wf) {
		const int line = wf * LP + l;
		const bool ok = line < nlines;
		if constexpr (DIR < 0) {
			const int cz = ok ? line / NC : 0, kx = ok ? line % NC : 0;
#pragma unroll
			for (int e = 0; e < E; e++) pre[e] = (ok && e < EL && t + T * e < nyd) ? Sx[sx_line(cz, kx, t + T * e, NC, ntl)] : make_float2(0.f, 0.f);
		}
		else {
			const float2* q = S + (size_t)(ok ? line : 0) * Sp;
#pragma unroll
			for (int e = 0; e < E; e++) pre[e] = ok ? q[t + T * e] : make_float2(0.f, 0.f);
		}
	};
	if ((int)blockIdx.x < nwork) fetch(blockIdx.x);
	for (int w = blockIdx.x; w < nwork; w += gridDim.x) {
		const int line = w * LP + l;
		C2<float> v[E];
#pragma unroll
		for (int e = 0; e < E; e++) v[e] = mk(pre[e].x, pre[e].y);
		if (w + (int)gridDim.x < nwork) fetch(w + gridDim.x);
		fft<PK, DIR, float>(v, lds + l * LL, 1, t, tw);
		if (line < nlines) {
			if constexpr (DIR < 0) {
				float2* p = S + (size_t)line * Sp;
#pragma unroll
				for (int e = 0; e < E; e++) st_nt(p + t + T * e, v[e].x, v[e].y);
			}
			else {
				const int cz = line / NC, kx = line % NC;
#pragma unroll
				for (int e = 0; e < EL; e++)
					if (t + T * e < nyd) st_nt(Sx + sx_line(cz, kx, t + T * e, NC, ntl), v[e].x, v[e].y);
			}
		}
	}
}

// ---------------------------------------------------------------------------------------
// P3 (3D): per (kx, ky) column, z-FFT of the nz data planes (zero padded to NZ), 6-term
// symmetric kernel multiply, z-IFFT, first nz planes written back.  Kernel symmetries
// (Boris/DemagKernel.cpp:85-168): diagonal even in y and z; Kxy odd in y; Kxz odd in z;
// Kyz odd in y and z.
// A workgroup takes CH columns ky = CH*b + c below N.y/2 and the CH mirror columns
// ky = N.y - CH*(b+1) + c at or above it, so both halves use kernel rows
// kyr in [CH*b, CH*b + CH]: that slab (all kz, 6 components, contiguous because the 3D
// kernel is stored kz fastest) is read from HBM once into LDS and serves the kz and N.z-kz
// mirrors of every column.
// FULL (periodic z, N.z = nz): all NZ planes hold data.  V: radix schedule of NZ.
template <int NZ, bool FULL = false, int V = 0>
__global__ void __launch_bounds__(kBlock) k_z(float2* __restrict__ S, const float* __restrict__ K,
	int nz, int NC, int Ny, int NYH, int Sp, int xmap, const PassBatch* __restrict__ bt)
{
	if (bt) { const PassBatch& b = bt[blockIdx.y]; S = b.S; K = b.K; nz = b.nz; NC = b.NC; Ny = b.Ny; NYH = b.NYH; Sp = b.Sp; }
	constexpr int PK = plan_key(NZ, V);
	constexpr int E = Plan<PK>::E, TZ = plan_T<PK>(), CY = kBlock / TZ, CH = CY / 2, NZH = NZ / 2 + 1, KR = CH + 1;
	constexpr int EL = FULL ? E : E / 2;
	constexpr int KS = 6 * KR * NZH, NKL = (KS + kBlock - 1) / kBlock;
	__shared__ C2<float> lds[3 * NZ * CY];
	__shared__ float ks[KS];
	const int col = threadIdx.x % CY, t = threadIdx.x / CY;
	const bool hi = col >= CH;
	const int c = hi ? col - CH : col;
	Twiddles<PK, float> tw; tw.init(t);
	const int half = Ny / 2, nwb = (half + CH - 1) / CH, nwork = NC * nwb;
	const size_t plane = (size_t)NC * Sp, comp = (size_t)nz * plane;
	const size_t kq = (size_t)NC * NYH * NZH;

	// Work item w: its columns' data (registers) and its kernel slab (registers, then LDS)
	// are fetched one item ahead, so loads overlap the transforms of the current item.
	float2 pd[3][EL];
	float pk[NKL];
	auto fetch = [&](int wf) {
		const int kx = wf / nwb, k0 = CH * (wf % nwb);
		const int ky = hi ? Ny - (k0 + CH) + c : k0 + c;
		const bool ok = hi ? (ky >= half && ky < Ny) : (ky < half);
		const float2* p = S + (size_t)kx * Sp + (ok ? ky : 0);
#pragma unroll
		for (int cc = 0; cc < 3; cc++)
#pragma unroll
			for (int e = 0; e < EL; e++) {
				const int z = t + TZ * e;
				pd[cc][e] = (ok && z < nz) ? p[cc * comp + z * plane] : make_float2(0.f, 0.f);
			}
		const int nrows = min(KR, NYH - k0);
		const float* kb = K + ((size_t)kx * NYH + k0) * NZH;
#pragma unroll
		for (int q = 0; q < NKL; q++) {
			const int i = threadIdx.x + kBlock * q;
			const int qq = i / (KR * NZH), rem = i - qq * (KR * NZH);
			pk[q] = (i < KS && rem / NZH < nrows) ? kb[qq * kq + rem] : 0.f;
		}
	};
	// xmap: XCD-contiguous order (each XCD sweeps consecutive ky chunks)
	const int G = gridDim.x, w0 = xmap ? xcd_chunk(blockIdx.x, G) : (int)blockIdx.x;
	if (w0 < nwork) fetch(w0);
	for (int w = w0; w < nwork; w += G) {
		const int kx = w / nwb, k0 = CH * (w % nwb);
		const int ky = hi ? Ny - (k0 + CH) + c : k0 + c;
		const bool ok = hi ? (ky >= half && ky < Ny) : (ky < half);
		const int lr = hi ? CH - c : c;                    // kernel row of this column, relative to k0
		C2<float> v[3][E];
#pragma unroll
		for (int cc = 0; cc < 3; cc++)
#pragma unroll
			for (int e = 0; e < E; e++) v[cc][e] = (e < EL) ? mk(pd[cc][e].x, pd[cc][e].y) : mk(0.f, 0.f);
		__syncthreads();                                   // the previous slab is no longer read
#pragma unroll
		for (int q = 0; q < NKL; q++) { const int i = threadIdx.x + kBlock * q; if (i < KS) ks[i] = pk[q]; }
		if (w + G < nwork) fetch(w + G);
		fft3<PK, -1, float>(v[0], v[1], v[2], lds + col, lds + NZ * CY + col, lds + 2 * NZ * CY + col, CY, t, tw);
		if constexpr (Plan<PK>::S == 1) __syncthreads();  // single-stage plans have no barrier inside
		const float sy = (ky <= half) ? 1.f : -1.f;
		const float* kr = ks + (ok ? lr : 0) * NZH;
#pragma unroll
		for (int e = 0; e < E; e++) {
			const int kz = t + TZ * e;
			const int kzr = kz <= NZ / 2 ? kz : NZ - kz;
			const float sz = kz <= NZ / 2 ? 1.f : -1.f;
			const float kxx = kr[kzr], kyy = kr[KR * NZH + kzr], kzz = kr[2 * KR * NZH + kzr];
			const float kxy = sy * kr[3 * KR * NZH + kzr], kxz = sz * kr[4 * KR * NZH + kzr], kyz = sy * sz * kr[5 * KR * NZH + kzr];
			C2<float> mx = v[0][e], my = v[1][e], mz = v[2][e];
			v[0][e] = mk(kxx * mx.x + kxy * my.x + kxz * mz.x, kxx * mx.y + kxy * my.y + kxz * mz.y);
			v[1][e] = mk(kxy * mx.x + kyy * my.x + kyz * mz.x, kxy * mx.y + kyy * my.y + kyz * mz.y);
			v[2][e] = mk(kxz * mx.x + kyz * my.x + kzz * mz.x, kxz * mx.y + kyz * my.y + kzz * mz.y);
		}
		fft3<PK, 1, float>(v[0], v[1], v[2], lds + col, lds + NZ * CY + col, lds + 2 * NZ * CY + col, CY, t, tw);
		if (ok) {
			float2* p = S + (size_t)kx * Sp + ky;
#pragma unroll
			for (int cc = 0; cc < 3; cc++)
#pragma unroll
				for (int e = 0; e < EL; e++) {
					const int z = t + TZ * e;
					if (z < nz) st_nt(p + cc * comp + z * plane, v[cc][e].x, v[cc][e].y);
				}
		}
	}
}

// workgroup size of P2' for plan key PK: 128 threads when the plan holds more than 8 values
// per thread (register pressure of three components), unless a line needs more threads
template <int PK> constexpr int y2d_block() { return (Plan<PK>::E > 8 && plan_T<PK>() <= 128) ? 128 : 256; }

// ---------------------------------------------------------------------------------------
// P2' (2D): per kx line of Sx (in place), y-FFT of 3 components, 2D kernel multiply (Kxy
// odd in y), y-IFFT.  A block takes LP consecutive lines, one per group of T lanes.
// nyd: data rows as in P2.  V: radix schedule of NY.
template <int NY, int BS, bool FULL = false, int V = 0>
__global__ void __launch_bounds__(BS) k_y2d(float2* __restrict__ Sx, const float* __restrict__ K, int NC, int NYH, int ntl,
	int nyd, const PassBatch* __restrict__ bt)
{
	if (bt) { const PassBatch& b = bt[blockIdx.y]; Sx = b.Sx; K = b.K; NC = b.NC; NYH = b.NYH; ntl = b.ntl; nyd = b.ny; }
	constexpr int PK = plan_key(NY, V);
	constexpr int E = Plan<PK>::E, T = plan_T<PK>(), LP = BS / T, LL = lds_len<NY>(), EL = FULL ? E : E / 2;
	__shared__ C2<float> lds[3 * LP * LL];
	const int t = threadIdx.x % T, l = threadIdx.x / T;
	Twiddles<PK, float> tw; tw.init(t);
	const size_t kq = (size_t)NC * NYH;
	const int nwork = (NC + LP - 1) / LP;
	for (int w = blockIdx.x; w < nwork; w += gridDim.x) {
		const int kx = w * LP + l;
		const bool ok = kx < NC;
		const int kxs = ok ? kx : 0;
		C2<float> v[3][E];
#pragma unroll
		for (int c = 0; c < 3; c++)
#pragma unroll
			for (int e = 0; e < E; e++) {
				const float2 a = (ok && e < EL && t + T * e < nyd) ? Sx[sx_line(c, kxs, t + T * e, NC, ntl)] : make_float2(0.f, 0.f);
				v[c][e] = mk(a.x, a.y);
			}
		C2<float>* b0 = lds + l * LL;
		fft3<PK, -1, float>(v[0], v[1], v[2], b0, b0 + LP * LL, b0 + 2 * LP * LL, 1, t, tw);
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
		fft3<PK, 1, float>(v[0], v[1], v[2], b0, b0 + LP * LL, b0 + 2 * LP * LL, 1, t, tw);
		if (ok) {
#pragma unroll
			for (int c = 0; c < 3; c++)
#pragma unroll
				for (int e = 0; e < EL; e++)
					if (t + T * e < nyd) Sx[sx_line(c, kx, t + T * e, NC, ntl)] = make_float2(v[c][e].x, v[c][e].y);
		}
	}
}

// ---------------------------------------------------------------------------------------
// Energy density  -mu0/(2n) sum M.H from the per-work-item partials of P5.
// (In1 + In2) / 2 per component, as the reference forms the averaged input
// (Boris/ConvolutionDataCUDA.cu:60-73)
__global__ void k_avg_in(const float* __restrict__ a, const float* __restrict__ b, float* __restrict__ out, size_t n3)
{
	for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < n3; q += (size_t)gridDim.x * blockDim.x)
		out[q] = (a[q] + b[q]) / 2;
}

// the field copied (clear) or added to both outputs (Boris/ConvolutionDataCUDA.cu:227-295)
__global__ void k_dup_out(const float* __restrict__ h, float* __restrict__ o1, float* __restrict__ o2, int clear, size_t n3)
{
	for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < n3; q += (size_t)gridDim.x * blockDim.x) {
		const float v = h[q];
		if (clear) { o1[q] = v; o2[q] = v; }
		else { o1[q] += v; o2[q] += v; }
	}
}

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
	int img[3];             // periodic images per axis (k_tensor_pbc)
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

// Periodic tensor (Boris/DemagTFunc_PBC.cpp:5-422), -N convention.  Along a periodic axis
// N equals the cell count and each offset o in 0..N-1 sums the open-boundary element at the
// image offsets o + k N, k = -p..p, in the reference's loop order (x images outermost); in
// 2D the z images sit at whole multiples of the cell thickness.  Off-diagonal terms take
// sign(offset) per axis with sign(0) = 0 (BorisLib/Funcs_Math.h:103).  Along an open axis
// offsets 0..n-1 are computed and mirrored (the mirror of a nonzero offset never coincides
// with another write).
__device__ __forceinline__ int isign0(int v) { return (v > 0) - (v < 0); }

__global__ void k_tensor_pbc(TensorArgs a, double* D, int is2d)
{
	const size_t NN = (size_t)a.Nx * a.Ny * a.Nz;
	const size_t total = (size_t)a.nx * a.ny * a.nz;
	const int d = kAsymDist, px = a.img[0], py = a.img[1], pz = a.img[2];
	const double hx = a.hx, hy = a.hy, hz = a.hz;
	for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < total; q += (size_t)gridDim.x * blockDim.x) {
		const int i0 = (int)(q % a.nx), j0 = (int)((q / a.nx) % a.ny), k0 = (int)(q / ((size_t)a.nx * a.ny));
		double vd[3] = { 0.0, 0.0, 0.0 }, vo[3] = { 0.0, 0.0, 0.0 };
		for (int ii = -px; ii <= px; ii++)
			for (int jj = -py; jj <= py; jj++)
				for (int kk = -pz; kk <= pz; kk++) {
					const int is = ii * a.Nx + i0, js = jj * a.Ny + j0, ks = is2d ? kk : kk * a.Nz + k0;
					const double si = isign0(is), sj = isign0(js), sk = isign0(ks);
					const int i = abs(is), j = abs(js), k = abs(ks);
					if (i >= d || j >= d || k >= d || i * i + j * j + k * k >= d * d) {
						vd[0] += -asym_diag(a.ad[0], i * hx, j * hy, k * hz);
						vd[1] += -asym_diag(a.ad[1], j * hy, i * hx, k * hz);
						vd[2] += -asym_diag(a.ad[2], k * hz, j * hy, i * hx);
						vo[0] += -asym_off(a.ao[0], i * hx, j * hy, k * hz) * (si * sj);
						if (!is2d) {
							vo[1] += -asym_off(a.ao[1], i * hx, k * hz, j * hy) * (si * sk);
							vo[2] += -asym_off(a.ao[2], j * hy, k * hz, i * hx) * (sj * sk);
						}
					}
					else {
						vd[0] += -stencil27(a.grid[0], a.gdim[0][0], a.gdim[0][1], i, j, k, hx, hy, hz);
						vd[1] += -stencil27(a.grid[1], a.gdim[1][0], a.gdim[1][1], j, i, k, hy, hx, hz);
						vd[2] += -stencil27(a.grid[2], a.gdim[2][0], a.gdim[2][1], k, j, i, hz, hy, hx);
						vo[0] += -stencil27(a.grid[3], a.gdim[3][0], a.gdim[3][1], i, j, k, hx, hy, hz) * (si * sj);
						if (!is2d) {
							vo[1] += -stencil27(a.grid[4], a.gdim[4][0], a.gdim[4][1], i, k, j, hx, hz, hy) * (si * sk);
							vo[2] += -stencil27(a.grid[5], a.gdim[5][0], a.gdim[5][1], j, k, i, hy, hz, hx) * (sj * sk);
						}
					}
				}
		const bool fx = !px && i0, fy = !py && j0, fz = !is2d && !pz && k0;
		for (int m = 0; m < 8; m++) {
			const int mx = m & 1, my = (m >> 1) & 1, mz = (m >> 2) & 1;
			if ((mx && !fx) || (my && !fy) || (mz && !fz)) continue;
			const int sx = mx ? -1 : 1, sy = my ? -1 : 1, sz = mz ? -1 : 1;
			const size_t idx = (size_t)((sx * i0 + a.Nx) % a.Nx) + (size_t)((sy * j0 + a.Ny) % a.Ny) * a.Nx
				+ (size_t)((sz * k0 + a.Nz) % a.Nz) * a.Nx * a.Ny;
			D[idx] = vd[0]; D[NN + idx] = vd[1]; D[2 * NN + idx] = vd[2];
			D[3 * NN + idx] = vo[0] * (sx * sy);
			if (!is2d) { D[4 * NN + idx] = vo[1] * (sx * sz); D[5 * NN + idx] = vo[2] * (sy * sz); }
		}
	}
}

// Real-input FFT of strided lines in FP64; keeps `keep` outputs, taking the real (part 0)
// or imaginary (part 1) part times `sign`.  Line l = a + na*b reads src[a*sa + b*sb + i*se].
template <int N, typename TO>
__global__ void __launch_bounds__(kBlock) k_line_fft(const double* __restrict__ src, TO* __restrict__ dst, int nlines, int na,
	long long sa, long long sb, long long se, long long da, long long db, long long de, int keep, int part, double sign)
{
	constexpr int E = Plan<N>::E, T = plan_T<N>(), LP = kBlock / T, LL = lds_len<N>();
	__shared__ C2<double> lds[LP * LL];
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
		fft<N, -1, double>(v, lds + l * LL, 1, t, tw);
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
#define BM_FOR_Z_SIZES(M) M(4) M(8) M(16) M(32) M(64) M(128) M(256)

bool x_supported(int n) { switch (n) {
#define C_(V_) case V_:
	BM_FOR_X_SIZES(C_)
#undef C_
	return true; default: return false; } }
bool y_supported(int n) { return x_supported(n); }
bool line_supported(int n) { return x_supported(n) || n == 1536 || n == 8192; }
bool z_supported(int n) { switch (n) {
#define C_(V_) case V_:
	BM_FOR_Z_SIZES(C_)
#undef C_
	return true; default: return false; } }

int grid_for(int nwork, int ncu, int per_cu) { return std::max(1, std::min(nwork, ncu * per_cu)); }
// grid of a pass whose blocks take XCD-contiguous work (xcd_chunk): a multiple of 8
int grid8(int nwork, int ncu, int per_cu) { return (grid_for(nwork, ncu, per_cu) + 7) / 8 * 8; }

// Row-tile height factor of the x passes (1 or 2 times TileY).  With the row-tiled Sx every
// tile is moved as whole runs either way, and the shorter tile (half the LDS, twice the
// resident blocks) measured faster for both passes at 512x512x32.  BM_XTILE_FWD /
// BM_XTILE_INV = 2 select the taller tile for tuning runs.
int xtile_factor(bool fwd)
{
	static int f[2] = { -1, -1 };
	int& v = f[fwd ? 1 : 0];
	if (v < 0) {
		const char* e = getenv(fwd ? "BM_XTILE_FWD" : "BM_XTILE_INV");
		v = (e && atoi(e) == 2) ? 2 : 1;
	}
	return v;
}

// full: periodic x (rows of 2 HX cells, no zero padding); only the plain transform exists
// for it (no fused RKDP54 advance, no taller tile).  nx: cells per row when below HX (the
// general-length variant; 0 = HX).
int launch_fwd_x(int HX, bool full, const float* M, float2* Sx, int ny, int nz, int ntl, int ncu, hipStream_t st,
	const float* E0 = nullptr, float* sM1 = nullptr, float* Mout = nullptr, float dT5 = 0.f, int nx = 0, int zoff = 0, int nzl = 0)
{
	// a periodic row holds 2 HX cells and is never of the general length
	const bool adv = E0 != nullptr, gen = !full && nx && nx != HX;
	if (full && adv) return fail(BM_ERR_STATE, "fused advance with a periodic x axis");
	if (gen && (full || adv)) return fail(BM_ERR_STATE, "general row length: no periodic x, no fused advance");
	const int f = (full || gen) ? 1 : xtile_factor(true);
	FwdArgs a{ M, Sx, ny, nz, ntl, 0, E0, sM1, Mout, dT5, gen ? nx : HX, zoff, nzl };
	switch (HX) {
#define C_(V_) case V_: { constexpr int TY1 = XTile<V_>::v; int TY = TY1 * f; a.ntiles = (ny + TY - 1) / TY; dim3 g(grid_for(nz * a.ntiles, ncu, 2)); \
		const FwdArgs* nb = nullptr; \
		if (full) hipLaunchKernelGGL((k_fwd_x<V_, TY1, false, true>), g, dim3(kBlock), 0, st, a, nb); \
		else if (gen) hipLaunchKernelGGL((k_fwd_x<V_, TY1, false, false, true>), g, dim3(kBlock), 0, st, a, nb); \
		else if (f == 1 && !adv) hipLaunchKernelGGL((k_fwd_x<V_, TY1, false>), g, dim3(kBlock), 0, st, a, nb); \
		else if (f == 1) hipLaunchKernelGGL((k_fwd_x<V_, TY1, true>), g, dim3(kBlock), 0, st, a, nb); \
		else if (!adv) hipLaunchKernelGGL((k_fwd_x<V_, 2 * TY1, false>), g, dim3(kBlock), 0, st, a, nb); \
		else hipLaunchKernelGGL((k_fwd_x<V_, 2 * TY1, true>), g, dim3(kBlock), 0, st, a, nb); return BM_OK; }
	BM_FOR_X_SIZES(C_)
#undef C_
	}
	return fail(BM_ERR_SIZE, "no x plan");
}

// full: periodic x, as in launch_fwd_x; gen: rows of nx < HX cells (any nx), as in launch_fwd_x
template <int HX, int TY1, bool FULL, bool GEN>
void launch_inv_x_var(int mode, dim3 g, hipStream_t st, const float2* Sx, float* out, int ny, int nz, int ntl, int nt, float scale,
	const float* M, double* epart, int nx, int zoff = 0, int nzl = 0)
{
	const PassBatch* nb = nullptr;
	if (mode == 0) hipLaunchKernelGGL((k_inv_x<HX, 0, false, TY1, FULL, GEN>), g, dim3(kBlock), 0, st, Sx, out, ny, nz, ntl, nt, scale, M, epart, nx, zoff, nzl, nb);
	else if (mode == 1 && !epart) hipLaunchKernelGGL((k_inv_x<HX, 1, false, TY1, FULL, GEN>), g, dim3(kBlock), 0, st, Sx, out, ny, nz, ntl, nt, scale, M, epart, nx, zoff, nzl, nb);
	else if (mode == 1) hipLaunchKernelGGL((k_inv_x<HX, 1, true, TY1, FULL, GEN>), g, dim3(kBlock), 0, st, Sx, out, ny, nz, ntl, nt, scale, M, epart, nx, zoff, nzl, nb);
	else if (!epart) hipLaunchKernelGGL((k_inv_x<HX, 2, false, TY1, FULL, GEN>), g, dim3(kBlock), 0, st, Sx, out, ny, nz, ntl, nt, scale, M, epart, nx, zoff, nzl, nb);
	else hipLaunchKernelGGL((k_inv_x<HX, 2, true, TY1, FULL, GEN>), g, dim3(kBlock), 0, st, Sx, out, ny, nz, ntl, nt, scale, M, epart, nx, zoff, nzl, nb);
}

int launch_inv_x(int HX, bool full, int mode, const float2* Sx, float* out, int ny, int nz, int ntl, float scale, int ncu, hipStream_t st,
	const float* M = nullptr, double* epart = nullptr, int nx = 0, int zoff = 0, int nzl = 0)
{
	const bool gen = !full && nx && nx != HX;
	const int f = (mode == 0 && !full && !gen) ? xtile_factor(false) : 1;
	switch (HX) {
#define C_(V_) case V_: { constexpr int TY1 = XTile<V_>::v; int TY = TY1 * f, nt = (ny + TY - 1) / TY; dim3 g(grid_for(3 * nz * nt, ncu, 2)); \
		const PassBatch* nb = nullptr; \
		if (full) launch_inv_x_var<V_, TY1, true, false>(mode, g, st, Sx, out, ny, nz, ntl, nt, scale, M, epart, HX, zoff, nzl); \
		else if (gen) launch_inv_x_var<V_, TY1, false, true>(mode, g, st, Sx, out, ny, nz, ntl, nt, scale, M, epart, nx, zoff, nzl); \
		else if (mode == 0 && f != 1) hipLaunchKernelGGL((k_inv_x<V_, 0, false, 2 * TY1>), g, dim3(kBlock), 0, st, Sx, out, ny, nz, ntl, nt, scale, M, epart, HX, 0, 0, nb); \
		else launch_inv_x_var<V_, TY1, false, false>(mode, g, st, Sx, out, ny, nz, ntl, nt, scale, M, epart, HX, zoff, nzl); return BM_OK; }
	BM_FOR_X_SIZES(C_)
#undef C_
	}
	return fail(BM_ERR_SIZE, "no x plan");
}

int inv_x_nwork(int HX, int ny, int nz)
{
	switch (HX) {
#define C_(V_) case V_: { int TY = XTile<V_>::v; return 3 * nz * ((ny + TY - 1) / TY); }
	BM_FOR_X_SIZES(C_)
#undef C_
	}
	return 0;
}

// z-pass work order switch for tuning runs: BM_Z_XCD=1 selects the XCD-contiguous order.
int z_xcd() { static int v = getenv("BM_Z_XCD") ? atoi(getenv("BM_Z_XCD")) : 0; return v; }
// Launch switches of the single-mesh y and z passes for tuning runs (0 or unset: defaults):
// BM_Y_RADIX / BM_Z_RADIX a radix schedule of the pass length (lengths without it take the
// default), BM_Y_PERCU / BM_Z_PERCU the workgroups per CU.
int env_int(const char* n) { const char* e = getenv(n); return e ? atoi(e) : 0; }
struct PassTune { int y_radix, y_percu, z_radix, z_percu; };
const PassTune& pass_tune()
{
	// negative values mean the defaults
	static const PassTune t{ std::max(0, env_int("BM_Y_RADIX")), std::max(0, env_int("BM_Y_PERCU")),
		std::max(0, env_int("BM_Z_RADIX")), std::max(0, env_int("BM_Z_PERCU")) };
	return t;
}
int launch_y_tuned(int NY, int dir, float2* S, float2* Sx, int nlines, int Sp, int NC, int ntl, int ny, int ncu, hipStream_t st);
int launch_z_tuned(int NZ, float2* S, const float* K, int nz, int NC, int Ny, int NYH, int Sp, int ncu, hipStream_t st);

template <int NY, bool FULL = false>
void launch_y_t(int dir, dim3 g, hipStream_t st, float2* S, float2* Sx, int nlines, int Sp, int NC, int ntl, int nyd, const PassBatch* bt)
{
	if (dir < 0) hipLaunchKernelGGL((k_y<NY, -1, FULL>), g, dim3(kBlock), 0, st, S, Sx, nlines, Sp, NC, ntl, nyd, bt);
	else hipLaunchKernelGGL((k_y<NY, 1, FULL>), g, dim3(kBlock), 0, st, S, Sx, nlines, Sp, NC, ntl, nyd, bt);
}

// full: periodic y (N.y = ny); ny: data rows of a line
int launch_y(int NY, bool full, int dir, float2* S, float2* Sx, int nlines, int Sp, int NC, int ntl, int ny, int ncu, hipStream_t st)
{
	if (!full && (pass_tune().y_radix || pass_tune().y_percu)) return launch_y_tuned(NY, dir, S, Sx, nlines, Sp, NC, ntl, ny, ncu, st);
	switch (NY) {
#define C_(V_) case V_: { int LP = kBlock / plan_T<V_>(); dim3 g(grid8((nlines + LP - 1) / LP, ncu, 8)); \
		if (full) launch_y_t<V_, true>(dir, g, st, S, Sx, nlines, Sp, NC, ntl, ny, nullptr); \
		else launch_y_t<V_>(dir, g, st, S, Sx, nlines, Sp, NC, ntl, ny, nullptr); return BM_OK; }
	BM_FOR_Y_SIZES(C_)
#undef C_
	}
	return fail(BM_ERR_SIZE, "no y plan");
}

// full: periodic z (N.z = nz)
int launch_z(int NZ, bool full, float2* S, const float* K, int nz, int NC, int Ny, int NYH, int Sp, int ncu, hipStream_t st)
{
	if (!full && (pass_tune().z_radix || pass_tune().z_percu)) return launch_z_tuned(NZ, S, K, nz, NC, Ny, NYH, Sp, ncu, st);
	switch (NZ) {
#define C_(V_) case V_: { int CH = kBlock / plan_T<V_>() / 2; int nw = NC * ((Ny / 2 + CH - 1) / CH); dim3 g(grid8(nw, ncu, 2)); \
		if (full) hipLaunchKernelGGL((k_z<V_, true>), g, dim3(kBlock), 0, st, S, K, nz, NC, Ny, NYH, Sp, z_xcd(), (const PassBatch*)nullptr); \
		else hipLaunchKernelGGL((k_z<V_>), g, dim3(kBlock), 0, st, S, K, nz, NC, Ny, NYH, Sp, z_xcd(), (const PassBatch*)nullptr); return BM_OK; }
	BM_FOR_Z_SIZES(C_)
#undef C_
	}
	return fail(BM_ERR_SIZE, "no z plan");
}

// full: periodic y (N.y = ny); ny: data rows of a line
int launch_y2d(int NY, bool full, float2* Sx, const float* K, int NC, int NYH, int ntl, int ny, int ncu, hipStream_t st)
{
	switch (NY) {
#define C_(V_) case V_: { constexpr int BS = y2d_block<V_>(); int LP = BS / plan_T<V_>(); dim3 g(grid8((NC + LP - 1) / LP, ncu, 8)); \
		if (full) hipLaunchKernelGGL((k_y2d<V_, BS, true>), g, dim3(BS), 0, st, Sx, K, NC, NYH, ntl, ny, (const PassBatch*)nullptr); \
		else hipLaunchKernelGGL((k_y2d<V_, BS>), g, dim3(BS), 0, st, Sx, K, NC, NYH, ntl, ny, (const PassBatch*)nullptr); return BM_OK; }
	BM_FOR_Y_SIZES(C_)
#undef C_
	}
	return fail(BM_ERR_SIZE, "no y plan");
}

// ---------------------------------------------------------------------------------------
// P2 + P3 + P4 fused for thin 3D meshes (N.z = NZ = 4 or 8, so nz <= NZ / 2): per kx, the
// 3 nz lines of Sx are y-transformed (one plane at a time, registers and the exchange buffer)
// into an LDS line buffer; every ky of those lines then has its nz values z-transformed
// (zero padded to NZ), multiplied by the 6-term symmetric kernel (signs as in P3) and
// inverse z-transformed (first nz planes kept) in place; the lines are inverse y-transformed
// back into Sx.  The full spectra S are never written: per evaluation the pass moves Sx twice
// and the kernel once, where P2 + P3 + P4 move Sx twice, S four times and the kernel once.
// A block takes LP consecutive kx, one per group of T lanes.  Open y and z only.
constexpr bool kYzyReg = true;
template <int NY, int NZ, int V, int BS>
__global__ void __launch_bounds__(BS) k_yzy(float2* __restrict__ Sx, const float* __restrict__ K, int nz, int NC, int NYH, int ntl,
	int nyd, const PassBatch* __restrict__ bt)
{
	if (bt) { const PassBatch& b = bt[blockIdx.y]; Sx = b.Sx; K = b.K; nz = b.nz; NC = b.NC; NYH = b.NYH; ntl = b.ntl; nyd = b.ny; }
	constexpr int PK = plan_key(NY, V);
	constexpr int E = Plan<PK>::E, T = plan_T<PK>(), LP = BS / T, LL = lds_len<NY>(), EL = E / 2, ZD = NZ / 2, NZH = NZ / 2 + 1;
	static_assert(LP * T == BS, "whole lines per block");
	// REG: the nz planes of the three components stay in registers (at most 24 complex values
	// per thread) and the z transforms run there, without the LDS line buffer
	constexpr bool REG = kYzyReg && ZD * 3 * E <= 24;
	__shared__ C2<float> lds[3 * LP * LL];            // exchange buffers of the y transforms
	__shared__ C2<float> lb[REG ? 1 : ZD * 3 * LP * NY];   // lines (z, c, l) of the y spectra
	const int t = threadIdx.x % T, l = threadIdx.x / T;
	Twiddles<PK, float> tw; tw.init(t);
	const size_t kq = (size_t)NC * NYH * NZH;
	const int nwork = (NC + LP - 1) / LP;
	C2<float>* b0 = lds + l * LL;
	auto line = [&](int z, int c, int ll) { return lb + ((z * 3 + c) * LP + ll) * NY; };
	for (int w = blockIdx.x; w < nwork; w += gridDim.x) {
		const int kx = w * LP + l;
		const bool ok = kx < NC;
		const int kxs = ok ? kx : 0;
		if constexpr (REG) {
			C2<float> v[ZD][3][E];
#pragma unroll
			for (int z = 0; z < ZD; z++) {
#pragma unroll
				for (int c = 0; c < 3; c++)
#pragma unroll
					for (int e = 0; e < E; e++) {
						const float2 a = (ok && z < nz && e < EL && t + T * e < nyd) ? Sx[sx_line(c * nz + z, kxs, t + T * e, NC, ntl)]
							: make_float2(0.f, 0.f);
						v[z][c][e] = mk(a.x, a.y);
					}
				if (z < nz) fft3<PK, -1, float>(v[z][0], v[z][1], v[z][2], b0, b0 + LP * LL, b0 + 2 * LP * LL, 1, t, tw);
			}
#pragma unroll
			for (int e = 0; e < E; e++) {
				const int ky = t + T * e;
				const int kyr = ky <= NY / 2 ? ky : NY - ky;
				const float sy = ky <= NY / 2 ? 1.f : -1.f;
				const float* kb = K + ((size_t)kxs * NYH + kyr) * NZH;
				C2<float> u[3][NZ];
#pragma unroll
				for (int c = 0; c < 3; c++) {
#pragma unroll
					for (int z = 0; z < NZ; z++) u[c][z] = z < ZD ? v[z][c][e] : mk(0.f, 0.f);
					Dft<NZ, -1, float>::run(u[c]);
				}
#pragma unroll
				for (int kz = 0; kz < NZ; kz++) {
					const int kzr = kz <= NZ / 2 ? kz : NZ - kz;
					const float sz = kz <= NZ / 2 ? 1.f : -1.f;
					const float kxx = kb[kzr], kyy = kb[kq + kzr], kzz = kb[2 * kq + kzr];
					const float kxy = sy * kb[3 * kq + kzr], kxz = sz * kb[4 * kq + kzr], kyz = sy * sz * kb[5 * kq + kzr];
					const C2<float> mx = u[0][kz], my = u[1][kz], mz = u[2][kz];
					u[0][kz] = mk(kxx * mx.x + kxy * my.x + kxz * mz.x, kxx * mx.y + kxy * my.y + kxz * mz.y);
					u[1][kz] = mk(kxy * mx.x + kyy * my.x + kyz * mz.x, kxy * mx.y + kyy * my.y + kyz * mz.y);
					u[2][kz] = mk(kxz * mx.x + kyz * my.x + kzz * mz.x, kxz * mx.y + kyz * my.y + kzz * mz.y);
				}
#pragma unroll
				for (int c = 0; c < 3; c++) {
					Dft<NZ, 1, float>::run(u[c]);
#pragma unroll
					for (int z = 0; z < ZD; z++) v[z][c][e] = u[c][z];
				}
			}
#pragma unroll
			for (int z = 0; z < ZD; z++) {
				if (z >= nz) continue;                   // uniform over the block
				fft3<PK, 1, float>(v[z][0], v[z][1], v[z][2], b0, b0 + LP * LL, b0 + 2 * LP * LL, 1, t, tw);
				if (ok) {
#pragma unroll
					for (int c = 0; c < 3; c++)
#pragma unroll
						for (int e = 0; e < EL; e++)
							if (t + T * e < nyd) Sx[sx_line(c * nz + z, kx, t + T * e, NC, ntl)] = make_float2(v[z][c][e].x, v[z][c][e].y);
				}
			}
			continue;
		}
		for (int z = 0; z < nz; z++) {                 // uniform over the block
			C2<float> v[3][E];
#pragma unroll
			for (int c = 0; c < 3; c++)
#pragma unroll
				for (int e = 0; e < E; e++) {
					const float2 a = (ok && e < EL && t + T * e < nyd) ? Sx[sx_line(c * nz + z, kxs, t + T * e, NC, ntl)] : make_float2(0.f, 0.f);
					v[c][e] = mk(a.x, a.y);
				}
			fft3<PK, -1, float>(v[0], v[1], v[2], b0, b0 + LP * LL, b0 + 2 * LP * LL, 1, t, tw);
#pragma unroll
			for (int c = 0; c < 3; c++)
#pragma unroll
				for (int e = 0; e < E; e++) line(z, c, l)[t + T * e] = v[c][e];
		}
		__syncthreads();
		for (int i = threadIdx.x; i < LP * NY; i += BS) {
			const int il = i / NY, ky = i % NY;
			const int kxi = w * LP + il < NC ? w * LP + il : 0;
			const int kyr = ky <= NY / 2 ? ky : NY - ky;
			const float sy = ky <= NY / 2 ? 1.f : -1.f;
			const float* kb = K + ((size_t)kxi * NYH + kyr) * NZH;
			C2<float> u[3][NZ];
#pragma unroll
			for (int c = 0; c < 3; c++) {
#pragma unroll
				for (int z = 0; z < NZ; z++) u[c][z] = z < nz ? line(z < ZD ? z : 0, c, il)[ky] : mk(0.f, 0.f);
				Dft<NZ, -1, float>::run(u[c]);
			}
#pragma unroll
			for (int kz = 0; kz < NZ; kz++) {
				const int kzr = kz <= NZ / 2 ? kz : NZ - kz;
				const float sz = kz <= NZ / 2 ? 1.f : -1.f;
				const float kxx = kb[kzr], kyy = kb[kq + kzr], kzz = kb[2 * kq + kzr];
				const float kxy = sy * kb[3 * kq + kzr], kxz = sz * kb[4 * kq + kzr], kyz = sy * sz * kb[5 * kq + kzr];
				const C2<float> mx = u[0][kz], my = u[1][kz], mz = u[2][kz];
				u[0][kz] = mk(kxx * mx.x + kxy * my.x + kxz * mz.x, kxx * mx.y + kxy * my.y + kxz * mz.y);
				u[1][kz] = mk(kxy * mx.x + kyy * my.x + kyz * mz.x, kxy * mx.y + kyy * my.y + kyz * mz.y);
				u[2][kz] = mk(kxz * mx.x + kyz * my.x + kzz * mz.x, kxz * mx.y + kyz * my.y + kzz * mz.y);
			}
#pragma unroll
			for (int c = 0; c < 3; c++) {
				Dft<NZ, 1, float>::run(u[c]);
#pragma unroll
				for (int z = 0; z < ZD; z++) if (z < nz) line(z, c, il)[ky] = u[c][z];
			}
		}
		__syncthreads();
		for (int z = 0; z < nz; z++) {                 // uniform over the block
			C2<float> v[3][E];
#pragma unroll
			for (int c = 0; c < 3; c++)
#pragma unroll
				for (int e = 0; e < E; e++) v[c][e] = line(z, c, l)[t + T * e];
			fft3<PK, 1, float>(v[0], v[1], v[2], b0, b0 + LP * LL, b0 + 2 * LP * LL, 1, t, tw);
			if (ok) {
#pragma unroll
				for (int c = 0; c < 3; c++)
#pragma unroll
					for (int e = 0; e < EL; e++)
						if (t + T * e < nyd) Sx[sx_line(c * nz + z, kx, t + T * e, NC, ntl)] = make_float2(v[c][e].x, v[c][e].y);
			}
		}
		__syncthreads();                               // the line buffer is rewritten by the next item
	}
}

// Schedule of NY for the fused thin-mesh pass with N.z = NZ: the radix-4 schedule (4 values
// per thread: the most lanes per line, so the most loads in flight per kx held in LDS) where
// the length has one, else one with at most 8 values per thread (NZ = 4 only); -1 if none.
template <int NY, int NZ> constexpr int yzy_variant()
{
	if constexpr (plan_variants(NY) > 2 && Plan<plan_key(NY, plan_variants(NY) > 2 ? 2 : 0)>::E <= 4) return 2;
	else if constexpr (NZ == 4 && Plan<plan_key(NY, 0)>::E <= 8) return 0;
	else if constexpr (NZ == 4 && plan_variants(NY) > 1 && Plan<plan_key(NY, plan_variants(NY) > 1 ? 1 : 0)>::E <= 8) return 1;
	else return -1;
}

template <int NY, int NZ> int launch_yzy_t(float2* Sx, const float* K, int nz, int NC, int NYH, int ntl, int ny, int ncu, hipStream_t st)
{
	constexpr int V = yzy_variant<NY, NZ>();
	if constexpr (V < 0) return fail(BM_ERR_SIZE, "no fused thin-mesh schedule for this length");
	else {
		constexpr int T = plan_T<plan_key(NY, V)>(), BS = T < 128 ? 128 : T, LP = BS / T;
		hipLaunchKernelGGL((k_yzy<NY, NZ, V, BS>), dim3(grid8((NC + LP - 1) / LP, ncu, 8)), dim3(BS), 0, st, Sx, K, nz, NC, NYH, ntl, ny,
			(const PassBatch*)nullptr);
		return BM_OK;
	}
}

// batched: the meshes of one (N.y, N.z) group, one per grid row
template <int NY, int NZ> int batch_yzy_t(LaunchShape sh, const PassBatch* h_tab, const PassBatch* d_tab, int nb, int ncu, hipStream_t st)
{
	constexpr int V = yzy_variant<NY, NZ>();
	if constexpr (V < 0) return fail(BM_ERR_SIZE, "no fused thin-mesh schedule for this length");
	else {
		constexpr int T = plan_T<plan_key(NY, V)>(), BS = T < 128 ? 128 : T, LP = BS / T;
		int mx = 1;
		for (int b = 0; b < nb; b++) mx = std::max(mx, (h_tab[b].NC + LP - 1) / LP);
		hipLaunchKernelGGL((k_yzy<NY, NZ, V, BS>), dim3(grid8(mx, ncu, sh.per_cu ? sh.per_cu : 8), nb), dim3(BS), 0, st,
			nullptr, nullptr, 0, 0, 0, 0, 0, d_tab);
		return BM_OK;
	}
}

#define BM_FOR_YZY_SIZES(M) M(16) M(32) M(64) M(128) M(256) M(512) M(1024) M(2048)

// the fused thin-mesh pass serves this mesh (open y and z, N.z 4 or 8, a power-of-two N.y
// with a schedule that fits).  By default only for N.z = 4 (two planes, held in registers with
// the radix-4 schedules) from 128 x 128 cells per plane: measured (RK4 steps, against the
// separate passes) 128x128x2 0.221 -> 0.207 ms, 256x256x2 0.256 -> 0.215, 512x512x2 0.364 ->
// 0.284, 1024x1024x2 0.994 -> 0.885 (N.y 2048: the LDS line buffer); for N.z = 8 (the LDS line
// buffer) 256x256x4 and 1024x512x4 within +-2 % either way, so not by default.
bool yzy_ok(const bm_demag_s* d)
{
	if (!d->yzy || d->is2d || d->full[1] || d->full[2] || (d->Nz != 4 && d->Nz != 8)) return false;
	if (d->yzy == 1 && (d->Nz != 4 || (long long)d->nx * d->ny < 128LL * 128)) return false;
	switch (d->Ny) {
#define C_(V_) case V_: return d->Nz == 4 ? yzy_variant<V_, 4>() >= 0 : yzy_variant<V_, 8>() >= 0;
	BM_FOR_YZY_SIZES(C_)
#undef C_
	}
	return false;
}

int launch_yzy(const bm_demag_s* d, hipStream_t st)
{
	switch (d->Ny) {
#define C_(V_) case V_: return d->Nz == 4 ? launch_yzy_t<V_, 4>(d->Sx, d->K, d->nz, d->NC, d->NYH, d->ntl, d->ny, d->ncu, st) \
	: launch_yzy_t<V_, 8>(d->Sx, d->K, d->nz, d->NC, d->NYH, d->ntl, d->ny, d->ncu, st);
	BM_FOR_YZY_SIZES(C_)
#undef C_
	}
	return fail(BM_ERR_SIZE, "no fused thin-mesh pass for this N.y");
}

// Real-input FFT of length 2M (the x transform of the tensor for N.x = 8192, beyond the one-
// block plans): the M-point complex FFT of z[k] = x[2k] + i x[2k+1], then the split
// X[k] = (Z[k] + conj Z[M-k]) / 2 - i W^k (Z[k] - conj Z[M-k]) / 2 and X[k+M] with -W^k,
// W = exp(-i pi / M), in FP64; arguments and outputs as k_line_fft's.
template <int M, typename TO>
__global__ void __launch_bounds__(kBlock) k_line_rfft2(const double* __restrict__ src, TO* __restrict__ dst, int nlines, int na,
	long long sa, long long sb, long long se, long long da, long long db, long long de, int keep, int part, double sign)
{
	constexpr int E = Plan<M>::E, T = plan_T<M>(), LP = kBlock / T, LL = lds_len<M>();
	static_assert(LP == 1, "one line per block");
	__shared__ C2<double> lds[LL];
	const int t = threadIdx.x;
	Twiddles<M, double> tw; tw.init(t);
	for (int line = blockIdx.x; line < nlines; line += gridDim.x) {
		const int a = line % na, b = line / na;
		const double* p = src + a * sa + b * sb;
		C2<double> v[E];
#pragma unroll
		for (int e = 0; e < E; e++) { const long long k = t + T * e; v[e] = mk(p[2 * k * se], p[(2 * k + 1) * se]); }
		fft<M, -1, double>(v, lds, 1, t, tw);
		__syncthreads();                              // the transform's last LDS reads are done
#pragma unroll
		for (int e = 0; e < E; e++) lds[lidx(t + T * e, 1)] = v[e];
		__syncthreads();
		TO* q = dst + a * da + b * db;
#pragma unroll
		for (int e = 0; e < E; e++) {
			const int k = t + T * e;
			const C2<double> zk = v[e], zm = conj(lds[lidx((M - k) % M, 1)]);
			const C2<double> fe = 0.5 * (zk + zm), fd = zk - zm;
			const C2<double> fo = mk(0.5 * fd.y, -0.5 * fd.x);        // -i (Z[k] - conj Z[M-k]) / 2
			double sn, cs;
			sincospi(-(double)k / M, &sn, &cs);
			const C2<double> wf = cmul(mk(cs, sn), fo);
			const C2<double> lo = fe + wf, hi = fe - wf;
			if (k < keep) q[k * de] = (TO)(sign * (part ? lo.y : lo.x));
			if (k + M < keep) q[(k + M) * de] = (TO)(sign * (part ? hi.y : hi.x));
		}
		__syncthreads();                              // before the next line's transform reuses the LDS
	}
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
	C_(1536)
#undef C_
	case 8192: hipLaunchKernelGGL((k_line_rfft2<4096, TO>), dim3(std::max(1, std::min(nlines, 8192))), dim3(kBlock), 0, st,
		src, dst, nlines, na, sa, sb, se, da, db, de, keep, part, sign); return BM_OK;
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
	// basis-function grids: the asymptotic distance along periodic axes (images reach it),
	// the cell count (at most that distance) along open ones
	const int dx = d->pbc[0] ? dd : std::min(nx, dd), dy = d->pbc[1] ? dd : std::min(ny, dd);
	const int dz = d->pbc[2] ? dd : std::min(nz, dd);
	const bool periodic = d->pbc[0] || d->pbc[1] || d->pbc[2];

	TensorArgs a;
	std::memset(&a, 0, sizeof(a));
	a.nx = nx; a.ny = ny; a.nz = nz; a.Nx = Nx; a.Ny = Ny; a.Nz = Nz; a.hx = hx; a.hy = hy; a.hz = hz;
	for (int q = 0; q < 3; q++) a.img[q] = d->pbc[q];
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
		const dim3 g((unsigned)std::min<size_t>((total + 127) / 128, 65535));
		if (periodic) hipLaunchKernelGGL(k_tensor_pbc, g, dim3(128), 0, st, a, D, d->is2d ? 1 : 0);
		else hipLaunchKernelGGL(k_tensor, g, dim3(128), 0, st, a, D, d->is2d ? 1 : 0);
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
			// z: line (kx, ky) -> K[kx][ky][kz] (kz fastest: the z pass stages kernel slabs contiguously)
			if (rc == BM_OK) rc = launch_line_fft<float>(Nz, T2, Kq, NC * NYH, NC, 1, NC, (long long)NC * NYH, (long long)NYH * NZH, NZH, 1, NZH, part[q][2], sgn, st);
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

// P1: the x-direction forward pass of M into the spectra.
int bm::demag_fwd_x(bm_demag_s* d, const float* M, hipStream_t st)
{
	prof_mark_begin(PK_FWD_X, st);
	int rc = launch_fwd_x(d->Nx / 2, d->full[0], M, d->Sx, d->ny, d->nz, d->ntl, d->ncu, st, nullptr, nullptr, nullptr, 0.f, d->nx);
	prof_mark_end(PK_FWD_X, st);
	return rc;
}

// P1 fused with the RKDP54 advance: sM1 = M, Mout = M + E0 dT/5, spectra of Mout.
int bm::demag_fwd_x_advance(bm_demag_s* d, const float* M, const float* E0, float* sM1, float* Mout, float dT5, hipStream_t st)
{
	prof_mark_begin(PK_FWD_ADV, st);
	int rc = launch_fwd_x(d->Nx / 2, d->full[0], M, d->Sx, d->ny, d->nz, d->ntl, d->ncu, st, E0, sM1, Mout, dT5);
	prof_mark_end(PK_FWD_ADV, st);
	return rc;
}

// The passes between the two x passes: y FFT, z FFT . K . z IFFT, y IFFT (3D) or the fused
// y FFT . K . y IFFT (2D).
int bm::demag_mid(bm_demag_s* d, hipStream_t st)
{
	const int nlines = 3 * d->nz * d->NC;
	int rc;
	if (d->is2d) {
		prof_mark_begin(PK_Y2D, st);
		rc = launch_y2d(d->Ny, d->full[1], d->Sx, d->K, d->NC, d->NYH, d->ntl, d->full[1] ? d->Ny : d->ny, d->ncu, st);
		prof_mark_end(PK_Y2D, st);
		return rc;
	}
	if (yzy_ok(d)) {
		// thin mesh: the y, z, y passes as one (timed as the z pass)
		prof_mark_begin(PK_Z, st);
		rc = launch_yzy(d, st);
		prof_mark_end(PK_Z, st);
		return rc;
	}
	prof_mark_begin(PK_Y_FWD, st);
	rc = launch_y(d->Ny, d->full[1], -1, d->S, d->Sx, nlines, d->Sp, d->NC, d->ntl, d->full[1] ? d->Ny : d->ny, d->ncu, st);
	prof_mark_end(PK_Y_FWD, st);
	if (rc) return rc;
	prof_mark_begin(PK_Z, st);
	rc = launch_z(d->Nz, d->full[2], d->S, d->K, d->nz, d->NC, d->Ny, d->NYH, d->Sp, d->ncu, st);
	prof_mark_end(PK_Z, st);
	if (rc) return rc;
	prof_mark_begin(PK_Y_INV, st);
	rc = launch_y(d->Ny, d->full[1], 1, d->S, d->Sx, nlines, d->Sp, d->NC, d->ntl, d->full[1] ? d->Ny : d->ny, d->ncu, st);
	prof_mark_end(PK_Y_INV, st);
	return rc;
}

// The convolution passes up to (not including) the final x C2R pass.
static int forward_and_multiply(bm_demag_s* d, const float* M, hipStream_t st)
{
	int rc = bm::demag_fwd_x(d, M, st);
	return rc ? rc : bm::demag_mid(d, st);
}

int bm::demag_planar(bm_demag_s* d, const float* M, float* H3n, hipStream_t st)
{
	int rc = forward_and_multiply(d, M, st);
	if (rc) return rc;
	const float scale = (float)(1.0 / ((double)d->Nx * d->Ny * d->Nz));
	prof_mark_begin(PK_INV_X, st);
	rc = launch_inv_x(d->Nx / 2, d->full[0], 0, d->Sx, H3n, d->ny, d->nz, d->ntl, scale, d->ncu, st, nullptr, nullptr, d->nx);
	prof_mark_end(PK_INV_X, st);
	return rc;
}

// The smallest x-pass length (2^a 3^b, 2..4096) at or above v, 0 if none.
int x_length_at_least(int v)
{
	int best = 0;
#define C_(V_) if (V_ >= v && (!best || V_ < best)) best = V_;
	BM_FOR_X_SIZES(C_)
#undef C_
	return best;
}

int bm::plan_length_at_least(int v) { return x_length_at_least(v); }

// The smallest y-pass length at or above v, 0 if none.
int y_length_at_least(int v) { return x_length_at_least(v); }

// Sizes (Boris/ConvolutionDataCUDA.cpp:24-100): N = n along a periodic axis, otherwise the
// reference takes N.x = 2 nx, N.y = 2 ny (any n, cuFFT plans every length) and N.z the next
// power of two >= 2 nz (N.z = 1 in 2D).  Here an open x or y axis takes the smallest length
// at or above 2 n that has a plan (N.x / 2 and N.y in the 2^a 3^b plans): any N >= 2 n - 1
// gives the same zero-padded (non-cyclic) convolution, so the field equals the reference's
// at N = 2 n up to FP32 rounding.  Meshes whose x rows are not exactly N.x / 2 cells long
// take the general-length x passes (and the unfused integrator evaluation).
static int demag_create(int nx, int ny, int nz, double hx, double hy, double hz, const int* pbc, bm_demag_t* out)
{
	if (!out || nx < 1 || ny < 1 || nz < 1 || !(hx > 0) || !(hy > 0) || !(hz > 0)) return fail(BM_ERR_ARG, "bm_demag_create: bad arguments");
	*out = nullptr;
	if (pbc && (pbc[0] < 0 || pbc[1] < 0 || pbc[2] < 0)) return fail(BM_ERR_ARG, "bm_demag_create_pbc: negative image count");
	const bool fx = pbc && pbc[0] > 0, fy = pbc && pbc[1] > 0, fz = nz > 1 && pbc && pbc[2] > 0;
	if (fx && nx % 2) return fail(BM_ERR_SIZE, "bm_demag_create_pbc: a periodic nx must be even");
	const int HX = fx ? nx / 2 : x_length_at_least(nx);
	const int Ny = fy ? ny : y_length_at_least(2 * ny);
	int Nz = 1;
	if (fz) Nz = nz;
	else if (nz > 1) { while (Nz < 2 * nz) Nz *= 2; }
	if (!HX || !Ny || !x_supported(HX) || !y_supported(Ny) || !line_supported(2 * HX) || (nz > 1 && !z_supported(Nz)))
		return fail(BM_ERR_SIZE, "bm_demag_create: no FFT plan for this mesh size (open axes: nx <= 4096, ny <= 2048, nz <= 128; "
			"periodic axes: N.x / 2 and N.y in the 2^a 3^b plans 2..4096, N.z a power of two 4..256, N = n)");
	bm_demag_s* d = new bm_demag_s();
	d->nx = nx; d->ny = ny; d->nz = nz; d->h[0] = hx; d->h[1] = hy; d->h[2] = hz;
	for (int q = 0; q < 3; q++) d->pbc[q] = pbc ? pbc[q] : 0;
	d->is2d = (nz == 1);
	d->full[0] = fx; d->full[1] = fy; d->full[2] = fz;
	d->Nx = 2 * HX; d->Ny = Ny; d->Nz = Nz;
	d->NC = d->Nx / 2 + 1; d->NYH = d->Ny / 2 + 1; d->NZH = d->Nz / 2 + 1;
	d->ncu = device_cu_count();
	if (const char* e = getenv("BM_ROW_WAVE")) d->row_wave = atoi(e) != 0;
	if (const char* e = getenv("BM_YZY")) d->yzy = std::max(0, std::min(2, atoi(e)));
	// line pitch of the full spectra (complex elements); BM_S_PAD adds padding for tuning experiments
	d->Sp = d->Ny;
	if (const char* e = getenv("BM_S_PAD")) d->Sp += std::max(0, atoi(e));
	d->ntl = sx_tiles(ny);
	const size_t sxbytes = (size_t)3 * nz * d->ntl * d->NC * kLT * sizeof(float2);
	const size_t sbytes = d->is2d ? 0 : (size_t)3 * nz * d->NC * d->Sp * sizeof(float2);
	d->kcount = d->is2d ? 4LL * d->NC * d->NYH : 6LL * d->NZH * d->NC * d->NYH;
	d->ecap = std::max(1, inv_x_nwork(d->Nx / 2, ny, nz));
	if (hipMalloc(&d->Sx, sxbytes) != hipSuccess || (sbytes && hipMalloc(&d->S, sbytes) != hipSuccess)
		|| hipMalloc(&d->K, d->kcount * sizeof(float)) != hipSuccess || hipMalloc(&d->ebuf, (size_t)d->ecap * sizeof(double)) != hipSuccess) {
		bm_demag_destroy(d);
		return fail(BM_ERR_NOMEM, "bm_demag_create: device memory");
	}
	d->bytes = (long long)(sxbytes + sbytes) + d->kcount * 4 + (long long)d->ecap * 8;
	hipStream_t st = nullptr;
	int rc = compute_kernels(d, st);
	// the allocation-free entry point's scratch address on this device, resolved at set-up
	float* sc = nullptr;
	if (!rc) rc = static_scratch(&sc);
	if (rc) { bm_demag_destroy(d); return rc; }
	*out = d;
	return BM_OK;
}

extern "C" int bm_demag_create(int nx, int ny, int nz, double hx, double hy, double hz, int flags, bm_demag_t* out)
{
	(void)flags;
	return demag_create(nx, ny, nz, hx, hy, hz, nullptr, out);
}

extern "C" int bm_demag_create_pbc(int nx, int ny, int nz, double hx, double hy, double hz, int pbc_x, int pbc_y, int pbc_z,
	int flags, bm_demag_t* out)
{
	(void)flags;
	const int pbc[3] = { pbc_x, pbc_y, pbc_z };
	return demag_create(nx, ny, nz, hx, hy, hz, pbc, out);
}

extern "C" int bm_demag_destroy(bm_demag_t d)
{
	if (!d) return BM_OK;
	if (d->Sx) (void)hipFree(d->Sx);
	if (d->S) (void)hipFree(d->S);
	if (d->K) (void)hipFree(d->K);
	if (d->ebuf) (void)hipFree(d->ebuf);
	if (d->avg) (void)hipFree(d->avg);
	if (d->hbuf) (void)hipFree(d->hbuf);
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
	const int HX = d->Nx / 2;
	int rc = forward_and_multiply(d, M, st);
	if (rc) return rc;
	const float scale = (float)(1.0 / ((double)d->Nx * d->Ny * d->Nz));
	if (energy) {
		const int np = inv_x_nwork(HX, d->ny, d->nz);
		if (np > d->ecap) return fail(BM_ERR_STATE, "energy partial buffer too small");
		prof_mark_begin(PK_INV_X, st);
		rc = launch_inv_x(HX, d->full[0], clear_out ? 1 : 2, d->Sx, Heff, d->ny, d->nz, d->ntl, scale, d->ncu, st, M, d->ebuf, d->nx);
		prof_mark_end(PK_INV_X, st);
		if (rc) return rc;
		size_t n = (size_t)d->nx * d->ny * d->nz;
		hipLaunchKernelGGL(k_energy_final, dim3(1), dim3(kBlock), 0, st, d->ebuf, np, n, energy);
	}
	else {
		prof_mark_begin(PK_INV_X, st);
		rc = launch_inv_x(HX, d->full[0], clear_out ? 1 : 2, d->Sx, Heff, d->ny, d->nz, d->ntl, scale, d->ncu, st, nullptr, nullptr, d->nx);
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

// ConvolutionCUDA::Convolute_AveragedInputs / _DuplicatedOutputs (Boris/ConvolutionCUDA.h:
// 225-247): the convolution of (M1 + M2) / 2, set or added to H1 and, when given, to H2; the
// energy is -mu0/2 <(M1 + M2)/2 . H> / n, the reference's -mu0 (In1 + In2) . H / (4 n).
extern "C" int bm_demag_convolute_averaged(bm_demag_t d, const float* M1, const float* M2, float* H1, float* H2,
	int clear_out, float* energy, void* stream)
{
	if (!d || !M1 || !M2 || !H1) return fail(BM_ERR_ARG, "bm_demag_convolute_averaged: null argument");
	hipStream_t st = (hipStream_t)stream;
	const size_t n3 = (size_t)d->nx * d->ny * d->nz * 3;
	if (!d->avg && hipMalloc(&d->avg, n3 * sizeof(float)) != hipSuccess) { d->avg = nullptr; return fail(BM_ERR_NOMEM, "averaged input buffer"); }
	if (H2 && !d->hbuf && hipMalloc(&d->hbuf, n3 * sizeof(float)) != hipSuccess) { d->hbuf = nullptr; return fail(BM_ERR_NOMEM, "field buffer"); }
	const int g = (int)std::min<size_t>((n3 + kBlock - 1) / kBlock, (size_t)d->ncu * 8);
	hipLaunchKernelGGL(k_avg_in, dim3(g), dim3(kBlock), 0, st, M1, M2, d->avg, n3);
	if (!H2) return bm_demag_convolute(d, d->avg, H1, clear_out, energy, stream);
	int rc = bm_demag_convolute(d, d->avg, d->hbuf, 1, energy, stream);
	if (rc) return rc;
	hipLaunchKernelGGL(k_dup_out, dim3(g), dim3(kBlock), 0, st, d->hbuf, H1, H2, clear_out ? 1 : 0, n3);
	BM_HIP(hipGetLastError());
	return BM_OK;
}

// ---------------------------------------------------------------------------------------
// Allocation-free averaged convolution.  The averaged input and, with a second output, the
// field buffer live in `scratch`, a file-scope device array each GPU holds until the process
// ends, instead of the handle's lazily allocated avg / hbuf: for callers that cannot call
// hipMalloc (graph capture, serving loops).  Meshes up to kStaticCells cells.  Every call on
// one GPU shares that GPU's array: calls that overlap on one GPU (other streams, other
// threads) race.  The array is fully written before it is read by each call (k_avg_in writes
// all 3n inputs, the set-mode convolution all 3n fields), so nothing needs clearing.
constexpr size_t kStaticCells = (size_t)1 << 20;
static __device__ __align__(256) float scratch[6 * kStaticCells];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
float* g_scratch_dev[kMaxDevices] = {};   // device address of `scratch` on each GPU
}

// Device address of `scratch` on the current device, resolved once per device.  Host code
// never uses the array's name as an address (on the host it is only a stand-in symbol).
int bm::static_scratch(float** out)
{
	int dev = 0;
	BM_HIP(hipGetDevice(&dev));
	if (dev < 0 || dev >= kMaxDevices) return fail(BM_ERR_STATE, "static scratch: device index out of range");
	std::lock_guard<std::mutex> lk(g_scratch_mu);
	if (!g_scratch_dev[dev]) {
		void* ptr = nullptr;
		const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
		if (e != hipSuccess || !ptr) return hip_fail(e != hipSuccess ? e : hipErrorInvalidSymbol, "hipGetSymbolAddress(scratch)");
		g_scratch_dev[dev] = (float*)ptr;
	}
	*out = g_scratch_dev[dev];
	return BM_OK;
}

extern "C" int bm_demag_convolute_averaged_static(bm_demag_t d, const float* M1, const float* M2, float* H1, float* H2,
	int clear_out, float* energy, void* stream)
{
	if (!d || !M1 || !M2 || !H1) return fail(BM_ERR_ARG, "bm_demag_convolute_averaged_static: null argument");
	const size_t n = (size_t)d->nx * d->ny * d->nz, n3 = 3 * n;
	if (n > kStaticCells)
		return fail(BM_ERR_SIZE, "bm_demag_convolute_averaged_static: mesh of " + std::to_string(n) + " cells exceeds the static scratch ("
			+ std::to_string(kStaticCells) + " cells)");
	float* sc = nullptr;
	int rc = static_scratch(&sc);
	if (rc) return rc;
	float* avg = sc;
	float* hb = sc + n3;
	hipStream_t st = (hipStream_t)stream;
	const int g = (int)std::min<size_t>((n3 + kBlock - 1) / kBlock, (size_t)d->ncu * 8);
	hipLaunchKernelGGL(k_avg_in, dim3(g), dim3(kBlock), 0, st, M1, M2, avg, n3);
	if (!H2) return bm_demag_convolute(d, avg, H1, clear_out, energy, stream);
	rc = bm_demag_convolute(d, avg, hb, 1, energy, stream);
	if (rc) return rc;
	hipLaunchKernelGGL(k_dup_out, dim3(g), dim3(kBlock), 0, st, hb, H1, H2, clear_out ? 1 : 0, n3);
	BM_HIP(hipGetLastError());
	return BM_OK;
}

extern "C" int bm_demag_copy_kernels(bm_demag_t d, float* host_out, long long count)
{
	if (!d || !host_out || count < d->kcount) return fail(BM_ERR_ARG, "bm_demag_copy_kernels: bad arguments");
	BM_HIP(hipMemcpy(host_out, d->K, d->kcount * sizeof(float), hipMemcpyDeviceToHost));
	return BM_OK;
}

// ---------------------------------------------------------------------------------------
// Pass launchers shared with the multilayer convolution (bm_mlayer.hip).
int bm::pass_fwd_x(int nx, const float* M, float2* Sx, int ny, int nplanes, int ntl, int ncu, hipStream_t st, bool full)
{
	prof_mark_begin(PK_FWD_X, st);
	int rc = launch_fwd_x(nx, full, M, Sx, ny, nplanes, ntl, ncu, st);
	prof_mark_end(PK_FWD_X, st);
	return rc;
}

int bm::pass_y(int Ny, int dir, float2* S, float2* Sx, int nlines, int Sp, int NC, int ntl, int ncu, hipStream_t st, bool full)
{
	const int kind = dir < 0 ? PK_Y_FWD : PK_Y_INV;
	prof_mark_begin(kind, st);
	int rc = launch_y(Ny, full, dir, S, Sx, nlines, Sp, NC, ntl, full ? Ny : Ny / 2, ncu, st);
	prof_mark_end(kind, st);
	return rc;
}

int bm::pass_inv_x(int nx, int mode, const float2* Sx, float* out, int ny, int nplanes, int ntl, float scale, int ncu, hipStream_t st,
	bool full)
{
	prof_mark_begin(PK_INV_X, st);
	int rc = launch_inv_x(nx, full, mode, Sx, out, ny, nplanes, ntl, scale, ncu, st);
	prof_mark_end(PK_INV_X, st);
	return rc;
}

int bm::setup_line_fft(int N, const double* src, double* dst, int nlines, int na, long long sa, long long sb, long long se,
	long long da, long long db, long long de, int keep, int part, double sign, hipStream_t st)
{
	return launch_line_fft<double>(N, src, dst, nlines, na, sa, sb, se, da, db, de, keep, part, sign, st);
}

int bm::setup_line_fft(int N, const double* src, float* dst, int nlines, int na, long long sa, long long sb, long long se,
	long long da, long long db, long long de, int keep, int part, double sign, hipStream_t st)
{
	return launch_line_fft<float>(N, src, dst, nlines, na, sa, sb, se, da, db, de, keep, part, sign, st);
}

int bm::pass_fwd_x_plane(int nx, const float* M, float2* Sx, int ny, int plane, int nzl, int ntl, int ncu, hipStream_t st, bool full)
{
	prof_mark_begin(PK_FWD_X, st);
	int rc = launch_fwd_x(nx, full, M, Sx, ny, 1, ntl, ncu, st, nullptr, nullptr, nullptr, 0.f, 0, plane, nzl);
	prof_mark_end(PK_FWD_X, st);
	return rc;
}

int bm::x_plane_parts(int nx, int ny) { return std::max(1, inv_x_nwork(nx, ny, 1)); }

int bm::pass_inv_x_plane(int nx, int mode, const float2* Sx, float* out, int ny, int plane, int nzl, int ntl, float scale, int ncu,
	hipStream_t st, const float* M, double* epart, float* energy, bool full)
{
	if (mode != 1 && mode != 2) return fail(BM_ERR_ARG, "plane x pass: AoS set or add");
	prof_mark_begin(PK_INV_X, st);
	int rc = launch_inv_x(nx, full, mode, Sx, out, ny, 1, ntl, scale, ncu, st, energy ? M : nullptr, energy ? epart : nullptr, 0, plane, nzl);
	prof_mark_end(PK_INV_X, st);
	if (rc || !energy) return rc;
	hipLaunchKernelGGL(k_energy_final, dim3(1), dim3(kBlock), 0, st, epart, inv_x_nwork(nx, ny, 1), (size_t)(full ? 2 * nx : nx) * ny, energy);
	return BM_OK;
}

bool bm::plan_2d_ok(int nx, int Ny) { return x_supported(nx) && x_supported(Ny) && line_supported(2 * nx); }

bool bm::plan_2d_pbc_ok(int nx, int ny, bool fx, bool fy)
{
	if (fx && nx % 2) return false;
	const int HX = fx ? nx / 2 : nx, Ny = fy ? ny : 2 * ny;
	return x_supported(HX) && line_supported(2 * HX) && y_supported(Ny);
}

// ---------------------------------------------------------------------------------------
// Batched pass launchers: nb meshes of one pass size; h_tab (host) sizes the grid, d_tab
// (device, same contents) is read by the kernels.
// f: 1 the base tile, 2 twice it, -2 half of it (where that still holds a whole row group)
int bm::x_tile_rows(int nx, bool fwd, int f)
{
	switch (nx) {
#define C_(V_) case V_: return f == -2 ? (x_half_tile<V_>() ? XTile<V_>::v / 2 : 0) : XTile<V_>::v * (f ? f : xtile_factor(fwd));
	BM_FOR_X_SIZES(C_)
#undef C_
	}
	return 0;
}

bool bm::x_half_tile_ok(int nx) { return x_tile_rows(nx, true, -2) > 0; }

// full: periodic x (rows of 2 HX cells, one row-tile height)
int bm::batch_fwd_x(int nx, bool full, LaunchShape sh, const FwdArgs* h_tab, const FwdArgs* d_tab, int nb, int ncu, hipStream_t st)
{
	int mx = 1;
	for (int b = 0; b < nb; b++) mx = std::max(mx, h_tab[b].nz * h_tab[b].ntiles);
	const int f = full ? 1 : sh.f ? sh.f : xtile_factor(true);
	const dim3 g(grid_for(mx, ncu, sh.per_cu ? sh.per_cu : 2), nb);
	FwdArgs a{};
	switch (nx) {
#define C_(V_) case V_: { constexpr int TY1 = XTile<V_>::v; \
		if (full) hipLaunchKernelGGL((k_fwd_x<V_, TY1, false, true>), g, dim3(kBlock), 0, st, a, d_tab); \
		else if (f == 1) hipLaunchKernelGGL((k_fwd_x<V_, TY1, false>), g, dim3(kBlock), 0, st, a, d_tab); \
		else if (f == -2) { if constexpr (x_half_tile<V_>()) hipLaunchKernelGGL((k_fwd_x<V_, TY1 / 2, false>), g, dim3(kBlock), 0, st, a, d_tab); \
			else return fail(BM_ERR_ARG, "no half tile for this row length"); } \
		else hipLaunchKernelGGL((k_fwd_x<V_, 2 * TY1, false>), g, dim3(kBlock), 0, st, a, d_tab); return BM_OK; }
	BM_FOR_X_SIZES(C_)
#undef C_
	}
	return fail(BM_ERR_SIZE, "no x plan");
}

int bm::batch_inv_x_planar(int nx, bool full, LaunchShape sh, const PassBatch* h_tab, const PassBatch* d_tab, int nb, int ncu, hipStream_t st)
{
	int mx = 1;
	for (int b = 0; b < nb; b++) mx = std::max(mx, 3 * h_tab[b].nz * h_tab[b].ntiles);
	const int f = full ? 1 : sh.f ? sh.f : xtile_factor(false);
	const dim3 g(grid_for(mx, ncu, sh.per_cu ? sh.per_cu : 2), nb);
	switch (nx) {
#define C_(V_) case V_: { constexpr int TY1 = XTile<V_>::v; \
		if (full) hipLaunchKernelGGL((k_inv_x<V_, 0, false, TY1, true>), g, dim3(kBlock), 0, st, nullptr, nullptr, 0, 0, 0, 0, 0.f, nullptr, nullptr, V_, 0, 0, d_tab); \
		else if (f == 1) hipLaunchKernelGGL((k_inv_x<V_, 0, false, TY1>), g, dim3(kBlock), 0, st, nullptr, nullptr, 0, 0, 0, 0, 0.f, nullptr, nullptr, V_, 0, 0, d_tab); \
		else if (f == -2) { if constexpr (x_half_tile<V_>()) hipLaunchKernelGGL((k_inv_x<V_, 0, false, TY1 / 2>), g, dim3(kBlock), 0, st, nullptr, nullptr, 0, 0, 0, 0, 0.f, nullptr, nullptr, V_, 0, 0, d_tab); \
			else return fail(BM_ERR_ARG, "no half tile for this row length"); } \
		else hipLaunchKernelGGL((k_inv_x<V_, 0, false, 2 * TY1>), g, dim3(kBlock), 0, st, nullptr, nullptr, 0, 0, 0, 0, 0.f, nullptr, nullptr, V_, 0, 0, d_tab); return BM_OK; }
	BM_FOR_X_SIZES(C_)
#undef C_
	}
	return fail(BM_ERR_SIZE, "no x plan");
}

// Batched passes with the radix schedule sh.radix (plan variant of the pass length).
namespace {
template <int NY, int V, bool FULL = false>
int batch_y_v(int dir, LaunchShape sh, const PassBatch* h_tab, const PassBatch* d_tab, int nb, int ncu, hipStream_t st)
{
	const int pc = sh.per_cu ? sh.per_cu : 8;
	constexpr int LP = kBlock / plan_T<plan_key(NY, V)>();
	int mx = 1;
	for (int b = 0; b < nb; b++) mx = std::max(mx, (h_tab[b].nlines + LP - 1) / LP);
	const dim3 g(grid8(mx, ncu, pc), nb);
	if (dir < 0) hipLaunchKernelGGL((k_y<NY, -1, FULL, V>), g, dim3(kBlock), 0, st, nullptr, nullptr, 0, 0, 0, 0, 0, d_tab);
	else hipLaunchKernelGGL((k_y<NY, 1, FULL, V>), g, dim3(kBlock), 0, st, nullptr, nullptr, 0, 0, 0, 0, 0, d_tab);
	return BM_OK;
}

template <int NZ, int V, bool FULL = false>
int batch_z_v(LaunchShape sh, const PassBatch* h_tab, const PassBatch* d_tab, int nb, int ncu, hipStream_t st)
{
	const int pc = sh.per_cu ? sh.per_cu : 2;
	constexpr int CH = kBlock / plan_T<plan_key(NZ, V)>() / 2;
	int mx = 1;
	for (int b = 0; b < nb; b++) mx = std::max(mx, h_tab[b].NC * ((h_tab[b].Ny / 2 + CH - 1) / CH));
	hipLaunchKernelGGL((k_z<NZ, FULL, V>), dim3(grid8(mx, ncu, pc), nb), dim3(kBlock), 0, st, nullptr, nullptr, 0, 0, 0, 0, 0, z_xcd(), d_tab);
	return BM_OK;
}

template <int NY, int V, bool FULL = false>
int batch_y2d_v(LaunchShape sh, const PassBatch* h_tab, const PassBatch* d_tab, int nb, int ncu, hipStream_t st)
{
	const int pc = sh.per_cu ? sh.per_cu : 8;
	constexpr int PK = plan_key(NY, V), BS = y2d_block<PK>(), LP = BS / plan_T<PK>();
	int mx = 1;
	for (int b = 0; b < nb; b++) mx = std::max(mx, (h_tab[b].NC + LP - 1) / LP);
	hipLaunchKernelGGL((k_y2d<NY, BS, FULL, V>), dim3(grid8(mx, ncu, pc), nb), dim3(BS), 0, st, nullptr, nullptr, 0, 0, 0, 0, d_tab);
	return BM_OK;
}

// schedule v of length N, or BM_ERR_ARG when N has no such schedule
template <int N, typename F> int with_variant(int v, F&& f)
{
	if (v == 0) return f(std::integral_constant<int, 0>());
	if constexpr (plan_variants(N) > 1) if (v == 1) return f(std::integral_constant<int, 1>());
	if constexpr (plan_variants(N) > 2) if (v == 2) return f(std::integral_constant<int, 2>());
	return fail(BM_ERR_ARG, "no such radix schedule for this length");
}

int launch_y_tuned(int NY, int dir, float2* S, float2* Sx, int nlines, int Sp, int NC, int ntl, int ny, int ncu, hipStream_t st)
{
	const int pc = pass_tune().y_percu ? pass_tune().y_percu : 8;
	switch (NY) {
#define C_(V_) case V_: return with_variant<V_>(pass_tune().y_radix < plan_variants(V_) ? pass_tune().y_radix : 0, [&](auto v) { \
		constexpr int V = decltype(v)::value, LP = kBlock / plan_T<plan_key(V_, V)>(); \
		const dim3 g(grid8((nlines + LP - 1) / LP, ncu, pc)); \
		if (dir < 0) hipLaunchKernelGGL((k_y<V_, -1, false, V>), g, dim3(kBlock), 0, st, S, Sx, nlines, Sp, NC, ntl, ny, (const PassBatch*)nullptr); \
		else hipLaunchKernelGGL((k_y<V_, 1, false, V>), g, dim3(kBlock), 0, st, S, Sx, nlines, Sp, NC, ntl, ny, (const PassBatch*)nullptr); \
		return (int)BM_OK; });
	BM_FOR_Y_SIZES(C_)
#undef C_
	}
	return fail(BM_ERR_SIZE, "no y plan");
}

int launch_z_tuned(int NZ, float2* S, const float* K, int nz, int NC, int Ny, int NYH, int Sp, int ncu, hipStream_t st)
{
	const int pc = pass_tune().z_percu ? pass_tune().z_percu : 2;
	switch (NZ) {
#define C_(V_) case V_: return with_variant<V_>(pass_tune().z_radix < plan_variants(V_) ? pass_tune().z_radix : 0, [&](auto v) { \
		constexpr int V = decltype(v)::value, CH = kBlock / plan_T<plan_key(V_, V)>() / 2; \
		const int nw = NC * ((Ny / 2 + CH - 1) / CH); \
		hipLaunchKernelGGL((k_z<V_, false, V>), dim3(grid8(nw, ncu, pc)), dim3(kBlock), 0, st, S, K, nz, NC, Ny, NYH, Sp, z_xcd(), \
			(const PassBatch*)nullptr); \
		return (int)BM_OK; });
	BM_FOR_Z_SIZES(C_)
#undef C_
	}
	return fail(BM_ERR_SIZE, "no z plan");
}
}

int bm::radix_variants(int n) { return plan_variants(n); }

const char* bm::radix_schedule(int n, int v)
{
	switch (plan_key(n, v)) {
#define S_(n_, v_) case plan_key(n_, v_): { static const std::string s = [] { std::string r; \
		for (int q = 0; q < Plan<plan_key(n_, v_)>::S; q++) r += (q ? "x" : "") + std::to_string(Plan<plan_key(n_, v_)>::R[q]); return r; }(); return s.c_str(); }
#define S3_(n_) S_(n_, 0) S_(n_, 1) S_(n_, 2)
#define S2_(n_) S_(n_, 0) S_(n_, 1)
	S2_(4) S3_(8) S3_(16) S3_(32) S3_(64) S2_(96) S3_(128) S2_(192) S3_(256) S2_(384) S3_(512) S2_(768) S3_(1024) S2_(2048)
	S_(2, 0) S_(6, 0) S_(12, 0) S_(24, 0) S_(48, 0) S_(1536, 0) S_(4096, 0)
#undef S_
#undef S2_
#undef S3_
	}
	return "";
}

// full: periodic axis (N = n), default radix schedule only
int bm::batch_y(int Ny, bool full, int dir, LaunchShape sh, const PassBatch* h_tab, const PassBatch* d_tab, int nb, int ncu, hipStream_t st)
{
	if (full && sh.radix) return fail(BM_ERR_ARG, "periodic passes take the default radix schedule");
	switch (Ny) {
#define C_(V_) case V_: if (full) return batch_y_v<V_, 0, true>(dir, sh, h_tab, d_tab, nb, ncu, st); \
		return with_variant<V_>(sh.radix, [&](auto v) { return batch_y_v<V_, decltype(v)::value>(dir, sh, h_tab, d_tab, nb, ncu, st); });
	BM_FOR_Y_SIZES(C_)
#undef C_
	}
	return fail(BM_ERR_SIZE, "no y plan");
}

int bm::batch_z(int Nz, bool full, LaunchShape sh, const PassBatch* h_tab, const PassBatch* d_tab, int nb, int ncu, hipStream_t st)
{
	if (full && sh.radix) return fail(BM_ERR_ARG, "periodic passes take the default radix schedule");
	switch (Nz) {
#define C_(V_) case V_: if (full) return batch_z_v<V_, 0, true>(sh, h_tab, d_tab, nb, ncu, st); \
		return with_variant<V_>(sh.radix, [&](auto v) { return batch_z_v<V_, decltype(v)::value>(sh, h_tab, d_tab, nb, ncu, st); });
	BM_FOR_Z_SIZES(C_)
#undef C_
	}
	return fail(BM_ERR_SIZE, "no z plan");
}

int bm::batch_y2d(int Ny, bool full, LaunchShape sh, const PassBatch* h_tab, const PassBatch* d_tab, int nb, int ncu, hipStream_t st)
{
	if (full && sh.radix) return fail(BM_ERR_ARG, "periodic passes take the default radix schedule");
	switch (Ny) {
#define C_(V_) case V_: if (full) return batch_y2d_v<V_, 0, true>(sh, h_tab, d_tab, nb, ncu, st); \
		return with_variant<V_>(sh.radix, [&](auto v) { return batch_y2d_v<V_, decltype(v)::value>(sh, h_tab, d_tab, nb, ncu, st); });
	BM_FOR_Y_SIZES(C_)
#undef C_
	}
	return fail(BM_ERR_SIZE, "no y plan");
}

bool bm::demag_yzy(const bm_demag_s* d) { return yzy_ok(d); }

int bm::batch_yzy(int Ny, int Nz, LaunchShape sh, const PassBatch* h_tab, const PassBatch* d_tab, int nb, int ncu, hipStream_t st)
{
	if (Nz != 4 && Nz != 8) return fail(BM_ERR_SIZE, "fused thin-mesh pass: N.z 4 or 8");
	switch (Ny) {
#define C_(V_) case V_: return Nz == 4 ? batch_yzy_t<V_, 4>(sh, h_tab, d_tab, nb, ncu, st) : batch_yzy_t<V_, 8>(sh, h_tab, d_tab, nb, ncu, st);
	BM_FOR_YZY_SIZES(C_)
#undef C_
	}
	return fail(BM_ERR_SIZE, "no fused thin-mesh pass for this N.y");
}
