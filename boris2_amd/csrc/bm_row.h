// Fused row kernel of the integrator: for TY rows of one z plane it
//   1. inverse-transforms the demag spectra of the three components along x (the last pass
//      of the convolution, H = IFFT(K . FFT(M)) / (Nx Ny Nz)), keeping H in registers;
//   2. runs the per-cell effective field + LLG + stage update (bm_stage.h) on those cells,
//      writing the stage's output magnetisation;
//   3. forward-transforms that output magnetisation along x into the same spectra rows (the
//      first pass of the next field evaluation's convolution).
// Each work item reads and then overwrites only its own rows y0..y0+TY-1 of the x-side
// spectra Sx (all kx, three components), so they are updated in place; neighbours for the exchange stencil are read from the
// stage's input magnetisation, which the kernel never writes.
//
// Compared with the separate passes (x-C2R writing H, stage reading H, x-R2C re-reading
// the new M) this removes 36 bytes of HBM traffic per cell per field evaluation and two
// launches.  Reference counterparts: ConvolutionCUDA::Convolute_3D's final inverse pass
// (Boris/ConvolutionDataCUDA.cu) + the evaluation kernels named in bm_stage.h + the next
// evaluation's forward pass.
#pragma once

#include "bm_internal.h"
#include "bm_stage.h"
#include "bm_xpass.h"

namespace bm {

template <int HX, int TY, int METHOD, int STEP, bool R2C, int BS = kXBlock>
__global__ void __launch_bounds__(BS) k_row(StageArgs a, float2* Sx, int ntl, int ntiles, float scale)
{
	constexpr int E = Plan<HX>::E, T = plan_T<HX>(), RP = BS / T, RPITCH = xrow_pitch<HX>(), NG = TY / RP, EH = E / 2;
	static_assert(NG * RP == TY, "tile height must be a multiple of the row group");
	__shared__ C2<float> tile[TY * RPITCH];
	XThread<HX> th; th.init();
	const int ny = a.ny, nz = a.nz;
	const int nwork = nz * ntiles;
	float lte_loc = 0.f;
	// gridDim.x is a multiple of 8: each sweep of gridDim.x items is split into 8 contiguous
	// parts, one per XCD, so neighbouring z planes of a y tile share an L2
	for (int r = 0; r * (int)gridDim.x < nwork; r++) {
		const int w = r * (int)gridDim.x + (int)xcd_chunk(blockIdx.x, gridDim.x);
		if (w >= nwork) continue;     // uniform over the block
		const int z = w % nz, y0 = (w / nz) * TY;
		// cells 2k and 2k+1 of this thread's rows: component c in h[g][e][c] (.x, .y)
		float2 h[NG][EH][3];
		XTileRegs<HX, TY, BS> pre;
		pre.load(Sx, 0, z, y0, ny, nz, ntl);
#pragma unroll 1
		for (int c = 0; c < 3; c++) {
			__syncthreads();
			pre.to_tile(tile);
			__syncthreads();
			if (c < 2) pre.load(Sx, c + 1, z, y0, ny, nz, ntl);
#pragma unroll
			for (int g = 0; g < NG; g++) {
				C2<float> v[E];
				c2r_row<HX>(th, tile + (g * RP + th.rl) * RPITCH, v);
#pragma unroll
				for (int e = 0; e < EH; e++) {
					// shift so that after the three components h[.][.][c] holds component c
					// (compile-time register indices only)
					h[g][e][0] = h[g][e][1]; h[g][e][1] = h[g][e][2];
					h[g][e][2] = make_float2(v[e].x * scale, v[e].y * scale);
				}
			}
		}
#pragma unroll
		for (int g = 0; g < NG; g++) {
			const int y = y0 + g * RP + th.rl;
			if (y < ny) {
#pragma unroll
				for (int e = 0; e < EH; e++) {
					const int x0 = 2 * (th.t + T * e);
					const size_t idx = ((size_t)z * ny + y) * HX + x0;
					const P6 Hd{ F3{ h[g][e][0].x, h[g][e][1].x, h[g][e][2].x }, F3{ h[g][e][0].y, h[g][e][1].y, h[g][e][2].y } };
					const P6 mn = stage_pair<METHOD, STEP>(a, idx, x0, y, z, ld6(a.Min, idx), Hd, lte_loc);
					h[g][e][0] = make_float2(mn.a.x, mn.b.x);
					h[g][e][1] = make_float2(mn.a.y, mn.b.y);
					h[g][e][2] = make_float2(mn.a.z, mn.b.z);
				}
			}
		}
		if (R2C) {
#pragma unroll 1
			for (int c = 0; c < 3; c++) {
#pragma unroll
				for (int g = 0; g < NG; g++) {
					C2<float> v[E];
#pragma unroll
					for (int e = 0; e < E; e++) {
						if (e < EH) {
							// component c is in slot 0 (rotated after each component)
							const float2 q = h[g][e][0];
							v[e] = mk(q.x, q.y);
							h[g][e][0] = h[g][e][1]; h[g][e][1] = h[g][e][2]; h[g][e][2] = q;
						}
						else v[e] = mk(0.f, 0.f);
					}
					r2c_row<HX>(th, tile + (g * RP + th.rl) * RPITCH, v);
				}
				__syncthreads();
				tile_to_S<HX, TY, BS>(tile, Sx, c, z, y0, ny, nz, ntl);
				__syncthreads();
			}
		}
	}
	if (stage_has_lte<METHOD, STEP>()) block_max_to<BS>(lte_loc, a.lte);
}

// Wave-independent variant for rows whose x transform is exactly one wavefront (HX = 512:
// 64 threads x 8 values).  The block takes the same work items as k_row (an 8-row tile of one
// z plane), but each of its four waves owns two rows of the tile end to end: it loads their
// spectra from Sx itself (both rows of one kx are adjacent in the row-tiled layout, so one
// 16-byte access per kx), transforms them through its own LDS rows with wavefront-scope
// ordering only, runs the stage on its cells and stores its two forward-transformed rows.
// No workgroup barrier separates the waves, so while one wave transforms (VALU and LDS)
// another can stream its stage arrays: the phases of the block's waves drift apart instead
// of alternating in lock step.  Same arithmetic in the same order as k_row (bit-identical).
template <int HX, int METHOD, int STEP, bool R2C>
__global__ void __launch_bounds__(kXBlock) k_roww(StageArgs a, float2* Sx, int ntl, int ntiles, float scale)
{
	constexpr int E = Plan<HX>::E, T = plan_T<HX>(), EH = E / 2, NC = HX + 1, RPITCH = xrow_pitch<HX>();
	constexpr int NW = kXBlock / 64, WY = kLT / NW, NQ = (NC + 63) / 64;
	static_assert(T == 64 && WY == 2, "one wavefront per row, two rows of the 8-row tile per wave");
	__shared__ C2<float> tile[NW * WY * RPITCH];
	XThread<HX> th; th.init();                 // th.t = lane, th.rl = wave
	const int lane = th.t;
	C2<float>* my = tile + th.rl * WY * RPITCH;
	const int ny = a.ny, nz = a.nz;
	const int nwork = nz * ntiles;
	float lte_loc = 0.f;
	for (int r = 0; r * (int)gridDim.x < nwork; r++) {
		const int w = r * (int)gridDim.x + (int)xcd_chunk(blockIdx.x, gridDim.x);
		if (w >= nwork) continue;                 // uniform over the block
		const int z = w % nz, y0 = (w / nz) * kLT + th.rl * WY;
		if (y0 >= ny) continue;                   // uniform over the wave
		const bool two = y0 + 1 < ny;             // the second row exists (ny may be odd)
		// both rows of kx = lane + 64 q as one 16-byte piece (y0 even: 16-byte aligned)
		float4 pre[NQ];
		auto load = [&](int c) {
			const float4* base = reinterpret_cast<const float4*>(Sx + sx_base(c, z, y0, nz, NC, ntl));
#pragma unroll
			for (int q = 0; q < NQ; q++) {
				const int kx = lane + 64 * q;
				float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
				if (kx < NC) v = base[(size_t)kx * (kLT / 2)];
				if (!two) { v.z = 0.f; v.w = 0.f; }
				pre[q] = v;
			}
		};
		float2 h[WY][EH][3];
		load(0);
#pragma unroll 1
		for (int c = 0; c < 3; c++) {
			xsync<true>();                        // the previous transforms' LDS reads are done
#pragma unroll
			for (int q = 0; q < NQ; q++) {
				const int kx = lane + 64 * q;
				if (kx < NC) { my[kx] = mk(pre[q].x, pre[q].y); my[RPITCH + kx] = mk(pre[q].z, pre[q].w); }
			}
			xsync<true>();
			if (c < 2) load(c + 1);
#pragma unroll
			for (int g = 0; g < WY; g++) {
				C2<float> v[E];
				c2r_row<HX, true>(th, my + g * RPITCH, v);
#pragma unroll
				for (int e = 0; e < EH; e++) {
					h[g][e][0] = h[g][e][1]; h[g][e][1] = h[g][e][2];
					h[g][e][2] = make_float2(v[e].x * scale, v[e].y * scale);
				}
			}
		}
#pragma unroll
		for (int g = 0; g < WY; g++) {
			const int y = y0 + g;
			if (y < ny) {
#pragma unroll
				for (int e = 0; e < EH; e++) {
					const int x0 = 2 * (lane + T * e);
					const size_t idx = ((size_t)z * ny + y) * HX + x0;
					const P6 Hd{ F3{ h[g][e][0].x, h[g][e][1].x, h[g][e][2].x }, F3{ h[g][e][0].y, h[g][e][1].y, h[g][e][2].y } };
					const P6 mn = stage_pair<METHOD, STEP>(a, idx, x0, y, z, ld6(a.Min, idx), Hd, lte_loc);
					h[g][e][0] = make_float2(mn.a.x, mn.b.x);
					h[g][e][1] = make_float2(mn.a.y, mn.b.y);
					h[g][e][2] = make_float2(mn.a.z, mn.b.z);
				}
			}
		}
		if (R2C) {
#pragma unroll 1
			for (int c = 0; c < 3; c++) {
#pragma unroll
				for (int g = 0; g < WY; g++) {
					C2<float> v[E];
#pragma unroll
					for (int e = 0; e < E; e++) {
						if (e < EH) {
							const float2 q = h[g][e][0];
							v[e] = mk(q.x, q.y);
							h[g][e][0] = h[g][e][1]; h[g][e][1] = h[g][e][2]; h[g][e][2] = q;
						}
						else v[e] = mk(0.f, 0.f);
					}
					r2c_row<HX, true>(th, my + g * RPITCH, v);
				}
				xsync<true>();
				float4* base = reinterpret_cast<float4*>(Sx + sx_base(c, z, y0, nz, NC, ntl));
#pragma unroll
				for (int q = 0; q < NQ; q++) {
					const int kx = lane + 64 * q;
					if (kx < NC) {
						const C2<float> u = my[kx], v = my[RPITCH + kx];
						if (two) base[(size_t)kx * (kLT / 2)] = make_float4(u.x, u.y, v.x, v.y);
						else reinterpret_cast<float2*>(base)[(size_t)kx * kLT] = make_float2(u.x, u.y);
					}
				}
			}
		}
	}
	if (stage_has_lte<METHOD, STEP>()) block_max_to<kXBlock>(lte_loc, a.lte);
}

// Grid of the row kernel: a multiple of 8 blocks (the XCD split needs it), about two per CU.
inline int row_grid(int nwork, int ncu)
{
	const int g = nwork < 2 * ncu ? nwork : 2 * ncu;
	return ((g + 7) / 8) * 8;
}

template <int METHOD, int STEP, bool R2C>
int launch_row_t(const bm_demag_s* d, const StageArgs& a, float scale, hipStream_t st)
{
	if (d->nx == 512 && d->row_wave) {
		constexpr int TY = kLT;
		const int nt = (d->ny + TY - 1) / TY;
		hipLaunchKernelGGL((k_roww<512, METHOD, STEP, R2C>), dim3(row_grid(d->nz * nt, d->ncu)), dim3(kXBlock), 0, st, a, d->Sx, d->ntl, nt, scale);
		return BM_OK;
	}
	switch (d->nx) {
#define C_(V_) case V_: { constexpr int TY = XTile<V_>::v; const int nt = (d->ny + TY - 1) / TY; \
		hipLaunchKernelGGL((k_row<V_, TY, METHOD, STEP, R2C>), dim3(row_grid(d->nz * nt, d->ncu)), dim3(kXBlock), 0, st, a, d->Sx, d->ntl, nt, scale); \
		return BM_OK; }
	BM_FOR_ROW_SIZES(C_)
#undef C_
	}
	return fail(BM_ERR_SIZE, "no x plan for the row kernel");
}

// per-method launchers (one translation unit each)
int launch_row_rkdp(const bm_demag_s* d, const StageArgs& a, int step, float scale, hipStream_t st);
int launch_row_rk4(const bm_demag_s* d, const StageArgs& a, int step, float scale, hipStream_t st);
int launch_row_heun(const bm_demag_s* d, const StageArgs& a, int mcode, int step, float scale, hipStream_t st);

} // namespace bm
