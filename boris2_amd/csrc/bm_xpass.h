// Device building blocks of the x-direction passes (real <-> half-spectrum FFT of mesh rows),
// shared by the stand-alone convolution passes and the fused row kernel of the integrator.
//
// A real row x[0..2HX) (the row of nx = HX cells zero padded to N.x = 2HX) is transformed as
// an HX-point complex FFT of z[k] = x[2k] + i x[2k+1] followed by the split into the
// N.x/2 + 1 = HX + 1 spectrum values; the inverse undoes the split and runs the HX-point
// inverse FFT, whose first HX/2 outputs are the nx real values.  Rows of one z plane are
// handled TY at a time; the spectrum rows sit in an LDS tile and move to and from the
// row-tiled x-side spectra Sx (layout below) in contiguous runs.
//
// Ownership: in a row group, thread t of row rl handles k = t + T*e, e < E/2, i.e. the cell
// pairs (2k, 2k+1) — the same cells in the inverse transform, in any per-cell work done in
// between, and in the forward transform.
#pragma once

#include "bm_fft.h"

namespace bm {

constexpr int kXBlock = 256;

// x-pass sizes HX = nx (the half-length complex FFT of a row zero padded to 2 nx); the fused
// row kernels of the integrator take them up to 2048, the stand-alone x passes up to 4096
#define BM_FOR_ROW_SIZES(M) M(2) M(4) M(6) M(8) M(12) M(16) M(24) M(32) M(48) M(64) M(96) M(128) M(192) M(256) M(384) M(512) M(768) M(1024) M(2048)
#define BM_FOR_X_SIZES(M) BM_FOR_ROW_SIZES(M) M(4096)
// y-pass lengths N.y (ny up to 2048)
#define BM_FOR_Y_SIZES(M) BM_FOR_X_SIZES(M)

// LDS tile row: HX+1 spectrum values, and room for the padded FFT exchange
template <int HX> constexpr int xrow_pitch() { return lds_len<HX>() > HX + 1 ? lds_len<HX>() : HX + 1; }

// rows per tile: at least one row group, at most 32 rows or about 66 KB of LDS
template <int HX> struct XTile {
	static constexpr int RP = kXBlock / plan_T<HX>();
	static constexpr int pw(int v) { int p = 1; while (p * 2 <= v) p *= 2; return p; }
	static constexpr int cap = pw(8400 / xrow_pitch<HX>()) < 32 ? pw(8400 / xrow_pitch<HX>()) : 32;
	static constexpr int v = RP > cap ? RP : cap;
};
// a tile of half the height still holds whole row groups (the batched passes can take it)
template <int HX> constexpr bool x_half_tile() { return XTile<HX>::v / 2 >= XTile<HX>::RP; }

// Per-thread constants of an x pass.
template <int HX> struct XThread {
	static constexpr int E = Plan<HX>::E, T = plan_T<HX>(), RP = kXBlock / T;
	int t, rl;
	Twiddles<HX, float> tw;
	C2<float> wk[E];   // W^k = exp(-i pi k / HX), k = t + T e

	__device__ __forceinline__ void init()
	{
		t = threadIdx.x % T; rl = threadIdx.x / T;
		tw.init(t);
#pragma unroll
		for (int e = 0; e < E; e++) { double s, c; sincospi(-(double)(t + T * e) / HX, &s, &c); wk[e] = mk((float)c, (float)s); }
	}
};

// Forward: v[e] = (x[2k], x[2k+1]) for e < E/2 and zero above (the zero-padded half).  On
// return LDS row lrow holds X[0..HX].  Every thread of the block must call it (barriers).
template <int HX, bool WAVE = false>
__device__ __forceinline__ void r2c_row(const XThread<HX>& th, C2<float>* lrow, C2<float>* v)
{
	constexpr int E = XThread<HX>::E, T = XThread<HX>::T;
	fft<HX, -1, float, WAVE>(v, lrow, 1, th.t, th.tw);
	xsync<WAVE>();
#pragma unroll
	for (int e = 0; e < E; e++) lrow[th.t + T * e] = v[e];
	xsync<WAVE>();
	C2<float> X[E];
#pragma unroll
	for (int e = 0; e < E; e++) {
		const int k = th.t + T * e;
		C2<float> zk = v[e], zm = conj(lrow[(HX - k) % HX]);
		C2<float> a = zk + zm, b = zk - zm;
		C2<float> wb = cmul(th.wk[e], b);                    // W^k (Z[k] - conj Z[HX-k])
		X[e] = mk(0.5f * (a.x + wb.y), 0.5f * (a.y - wb.x)); // a/2 - i wb/2
	}
	xsync<WAVE>();
#pragma unroll
	for (int e = 0; e < E; e++) lrow[th.t + T * e] = X[e];
	if (th.t == 0) lrow[HX] = mk(v[0].x - v[0].y, 0.f);
}

// Inverse: LDS row lrow holds X[0..HX].  On return v[e], e < E/2, holds the unscaled real
// pair (x[2n], x[2n+1]), n = t + T e (N.x times the true values).
template <int HX, bool WAVE = false>
__device__ __forceinline__ void c2r_row(const XThread<HX>& th, C2<float>* lrow, C2<float>* v)
{
	constexpr int E = XThread<HX>::E, T = XThread<HX>::T;
#pragma unroll
	for (int e = 0; e < E; e++) {
		const int k = th.t + T * e;
		C2<float> xk = lrow[k], xm = conj(lrow[HX - k]);
		C2<float> a = xk + xm, b = xk - xm;
		C2<float> wb = cmulc(b, th.wk[e]);                   // W^-k (X[k] - conj X[HX-k])
		v[e] = mk(a.x - wb.y, a.y + wb.x);                    // a + i wb
	}
	fft<HX, 1, float, WAVE>(v, lrow, 1, th.t, th.tw);         // first LDS write waits on a barrier
}

// x-side spectra layout.  The x passes and the y passes exchange the spectra of the ny data
// rows through Sx[c][z][t][kx][kLT]: row y of line (c, z, kx) sits in row tile t = y / kLT,
// slot y % kLT.  One row tile of a plane is a contiguous run of NC * kLT complex values, so
// an x pass over kLT rows (or a multiple) moves whole runs, and the y passes read kLT-row
// pieces of several consecutive kx lines at once (also contiguous).  ntl = ceil(ny / kLT)
// tiles per plane; slots of rows >= ny are padding and never read.
constexpr int kLT = 8;

__host__ __device__ __forceinline__ int sx_tiles(int ny) { return (ny + kLT - 1) / kLT; }
// index of (c, z, kx = 0, y0); for TY < kLT the tile's rows y0..y0+TY-1 share one row tile
__host__ __device__ __forceinline__ size_t sx_base(int c, int z, int y0, int nz, int NC, int ntl)
{
	return (((size_t)c * nz + z) * ntl + y0 / kLT) * (size_t)NC * kLT + y0 % kLT;
}
// index of row y of line (plane cz = c * nz + z, kx)
__host__ __device__ __forceinline__ size_t sx_line(int cz, int kx, int y, int NC, int ntl)
{
	return ((size_t)cz * ntl + y / kLT) * (size_t)NC * kLT + (size_t)kx * kLT + y % kLT;
}

// Element idx of a TY-row tile (TY a power of two) with NC values per row: its kx, row yl
// (relative to y0) and offset from sx_base.  Consecutive idx are consecutive addresses
// (TY >= kLT) or runs of TY (TY < kLT).
template <int TY, int NC>
__device__ __forceinline__ void sx_tile_elem(int idx, int& kx, int& yl, int& off)
{
	if constexpr (TY >= kLT) {
		const int q = idx / kLT;
		kx = q % NC; yl = (q / NC) * kLT + idx % kLT; off = idx;
	}
	else { kx = idx / TY; yl = idx % TY; off = kx * kLT + yl; }
}

// Copy of a tile of TY spectrum rows (component c, plane z, rows y0..) from LDS to Sx by a
// block of BS threads.
template <int HX, int TY, int BS = kXBlock>
__device__ __forceinline__ void tile_to_S(const C2<float>* tile, float2* Sx, int c, int z, int y0, int ny, int nz, int ntl)
{
	constexpr int NC = HX + 1, RPITCH = xrow_pitch<HX>();
	float2* base = Sx + sx_base(c, z, y0, nz, NC, ntl);
	for (int idx = threadIdx.x; idx < TY * NC; idx += BS) {
		int kx, yl, off;
		sx_tile_elem<TY, NC>(idx, kx, yl, off);
		if (y0 + yl < ny) {
			C2<float> a = tile[yl * RPITCH + kx];
			base[off] = make_float2(a.x, a.y);
		}
	}
}

// Register staging of a tile of TY spectrum rows read from Sx, so the next tile's loads can
// be in flight while the current one is transformed.
template <int HX, int TY, int BS = kXBlock> struct XTileRegs {
	static constexpr int NC = HX + 1, N = (TY * NC + BS - 1) / BS;
	float2 r[N];

	__device__ __forceinline__ void load(const float2* Sx, int c, int z, int y0, int ny, int nz, int ntl)
	{
		const float2* base = Sx + sx_base(c, z, y0, nz, NC, ntl);
#pragma unroll
		for (int q = 0; q < N; q++) {
			const int idx = threadIdx.x + BS * q;
			int kx, yl, off;
			sx_tile_elem<TY, NC>(idx, kx, yl, off);
			r[q] = (idx < TY * NC && y0 + yl < ny) ? base[off] : make_float2(0.f, 0.f);
		}
	}
	__device__ __forceinline__ void to_tile(C2<float>* tile) const
	{
		constexpr int RPITCH = xrow_pitch<HX>();
#pragma unroll
		for (int q = 0; q < N; q++) {
			const int idx = threadIdx.x + BS * q;
			int kx, yl, off;
			sx_tile_elem<TY, NC>(idx, kx, yl, off);
			if (idx < TY * NC) tile[yl * RPITCH + kx] = mk(r[q].x, r[q].y);
		}
	}
};

} // namespace bm
