// Multilayered demag of stacked 2D layers on MI355X: L layers of the same nx x ny cells
// (cell sizes hx, hy, hz[l]), each layer's bottom face at z0[l];
//   H_d = IFFT( sum_s K_ds . FFT(M_s) ) / (Nx Ny),  N = (2 nx, 2 ny).
//
// Reference behaviour: SDemagCUDA::UpdateField with multilayered convolution
// (Boris/SDemagCUDA.cpp:270) over DemagKernelCollectionCUDA (kernels
// Boris/DemagKernelCollectionCUDA_Calc*.cpp, product Boris/DemagKernelCollectionCUDA_Mult.cu:610)
// for 2D layers related by z shifts, the configuration SDemag builds for stacked layers of
// equal extent (Boris/SDemag.cpp:173-300).  Kernels: the layer's own 2D kernel for d == s;
// otherwise the z-shifted tensor of Boris/DemagTFunc_Shifted.cpp (equal cells) or
// Boris/DemagTFunc_Irregular.cpp (thicknesses differ) turned into six real kernels
// (xx, yy, zz, xy real; xz, yz the coefficients of i), Boris/DemagKernelCollection_Calc.cpp:288-420.
//
// Pipeline per evaluation (all layers at once, layers in place of z planes of the spectra
// S[c][l][kx][ky]): x R2C pass, y FFT pass, one pass that forms every destination's
// kernel sum from all layers' spectra at a (kx, ky) point in registers, y IFFT pass, x C2R
// pass.  The x and y passes are the single-mesh convolution's own.
#include "bm_internal.h"
#include "bm_newell.h"
#include "bm_xpass.h"
#include "bm_stage.h"   // kMU0

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

namespace {

using namespace bm;

// Layers per handle: up to kRegLayers the product keeps every layer's spectra in registers
// and works in place; above it a general product writes to a second spectra array.
constexpr int kRegLayers = 16;
constexpr int kMaxLayers = 64;

// Sparse transfer between the layers' own cells and the common planes (CSR rows: row r sums
// val[k] * src[col[k]] over k in [ptr[r], ptr[r + 1]), three components).  Transfer in: one
// row per common cell of every layer plane (l * plane + c), columns the layer cells of the
// concatenated layer arrays; transfer out: one row per layer cell, columns common cells of the
// layer's plane (relative to it).
struct Csr { const int* ptr; const int* col; const double* val; };

// rows r0 .. r0+nr-1 of a transfer-in table into dst rows 0.., sources offset by msub
__global__ void k_ml_transfer_in(Csr t, const float* __restrict__ M, float* __restrict__ dst, long long r0, long long nr, long long msub)
{
	for (long long q = blockIdx.x * (long long)blockDim.x + threadIdx.x; q < nr; q += (long long)gridDim.x * blockDim.x) {
		double a0 = 0, a1 = 0, a2 = 0;
		for (int k = t.ptr[r0 + q]; k < t.ptr[r0 + q + 1]; k++) {
			const size_t c = (size_t)(t.col[k] - msub);
			const double w = t.val[k];
			a0 += w * M[3 * c]; a1 += w * M[3 * c + 1]; a2 += w * M[3 * c + 2];
		}
		float* o = dst + 3 * q;
		o[0] = (float)a0; o[1] = (float)a1; o[2] = (float)a2;
	}
}

// rows r0 .. r0+nr-1 of a transfer-out table (layer cells) from the common plane P, set
// (mode 1) or added (mode 2) into H rows 0..; with M (the same layer cells) each block also
// leaves its partial sum of M . (the transferred field) in epart[blockIdx.x] (fixed order).
__global__ void __launch_bounds__(256) k_ml_transfer_out(Csr t, const float* __restrict__ P, float* __restrict__ H, long long r0, long long nr,
	int mode, const float* __restrict__ M, double* __restrict__ epart)
{
	__shared__ double red[256];
	double es = 0;
	for (long long q = blockIdx.x * (long long)blockDim.x + threadIdx.x; q < nr; q += (long long)gridDim.x * blockDim.x) {
		double a0 = 0, a1 = 0, a2 = 0;
		for (int k = t.ptr[r0 + q]; k < t.ptr[r0 + q + 1]; k++) {
			const size_t c = (size_t)t.col[k];
			const double w = t.val[k];
			a0 += w * P[3 * c]; a1 += w * P[3 * c + 1]; a2 += w * P[3 * c + 2];
		}
		float* o = H + 3 * q;
		const float h0 = (float)a0, h1 = (float)a1, h2 = (float)a2;
		if (mode == 2) { o[0] += h0; o[1] += h1; o[2] += h2; }
		else { o[0] = h0; o[1] = h1; o[2] = h2; }
		if (M) es += (double)M[3 * q] * h0 + (double)M[3 * q + 1] * h1 + (double)M[3 * q + 2] * h2;
	}
	if (!M) return;
	red[threadIdx.x] = es;
	__syncthreads();
	for (int o = 128; o > 0; o >>= 1) { if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o]; __syncthreads(); }
	if (threadIdx.x == 0) epart[blockIdx.x] = red[0];
}

// energy density -mu0/(2 n) sum of the partials, summed in order
__global__ void k_ml_energy(const double* part, int np, long long n, float* energy)
{
	if (threadIdx.x || blockIdx.x) return;
	double s = 0;
	for (int q = 0; q < np; q++) s += part[q];
	*energy = (float)(-kMU0 / 2.0 * s / (double)n);
}



// Position of grid index i (-1, 0, +1, ...) along one axis: i*h + o, or a table for -1, 0, +1.
// grid axis: linear i h + o, a table of three plane positions (i = -1, 0, 1), or centred
// (cen): grid index i holds offset i + lo + 1 at (i + lo + 1) h + o
struct Axis { double h, o, p[3]; int table, lo; };
__host__ __device__ inline double axis_pos(const Axis& a, int i)
{
	return a.table == 1 ? a.p[i + 1] : a.table == 2 ? (i + a.lo + 1) * a.h + a.o : i * a.h + a.o;
}
Axis lin(double h, double o) { Axis a; a.h = h; a.o = o; a.p[0] = a.p[1] = a.p[2] = 0; a.table = 0; a.lo = 0; return a; }
Axis tab(double m, double z, double p) { Axis a; a.h = a.o = 0; a.p[0] = m; a.p[1] = z; a.p[2] = p; a.table = 1; a.lo = 0; return a; }
Axis cen(double h, double o, int lo) { Axis a = lin(h, o); a.table = 2; a.lo = lo; return a; }

// Basis-function grid over i in [-1, dx], j in [-1, dy], k in [-1, dz] (stored with +1, x
// fastest); dz = -1 gives the single plane k = -1.
__global__ void k_fill_grid_at(double* g, int which, int dx, int dy, int dz, Axis X, Axis Y, Axis Z)
{
	const int gx = dx + 2, gy = dy + 2, gz = dz + 2;
	const size_t total = (size_t)gx * gy * gz;
	for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < total; q += (size_t)gridDim.x * blockDim.x) {
		const int i = (int)(q % gx) - 1, j = (int)((q / gx) % gy) - 1, k = (int)(q / ((size_t)gx * gy)) - 1;
		const double x = axis_pos(X, i), y = axis_pos(Y, j), z = axis_pos(Z, k);
		g[q] = which ? newell_g(x, y, z) : newell_f(x, y, z);
	}
}

// 36-term second differences for cells whose source and destination thicknesses differ
// (Boris/DemagTFunc_L.cpp:596-830): g holds three planes, del one extra plane; tau is the
// destination cell volume.  Term order as the reference's (compensated sum).
struct G3 {
	const double* v; int gx, gy;
	__device__ double operator()(int a, int b, int c) const { return v[(size_t)a + (size_t)b * gx + (size_t)c * gx * gy]; }
};

// xx, yy, xy: the plane axis is the third; evaluated at (i, j, 0)
__device__ double L36_plane3(const G3& G, const G3& Dl, int i, int j, double tau)
{
	const int k = 1; i++; j++;
	auto D = [&](int a, int b) { return Dl(a, b, 0); };
	double t[36] = {
		+4 * G(i, j, k), +4 * D(i, j),
		-2 * G(i + 1, j, k), -2 * D(i + 1, j), -2 * G(i - 1, j, k), -2 * D(i - 1, j),
		-2 * G(i, j + 1, k), -2 * D(i, j + 1), -2 * G(i, j - 1, k), -2 * D(i, j - 1),
		-4 * G(i, j, k + 1), -4 * G(i, j, k - 1),
		+G(i + 1, j + 1, k), +D(i + 1, j + 1), +G(i + 1, j - 1, k), +D(i + 1, j - 1),
		+G(i - 1, j + 1, k), +D(i - 1, j + 1), +G(i - 1, j - 1, k), +D(i - 1, j - 1),
		+2 * G(i + 1, j, k + 1), +2 * G(i + 1, j, k - 1), +2 * G(i - 1, j, k + 1), +2 * G(i - 1, j, k - 1),
		+2 * G(i, j + 1, k + 1), +2 * G(i, j + 1, k - 1), +2 * G(i, j - 1, k + 1), +2 * G(i, j - 1, k - 1),
		-G(i - 1, j - 1, k - 1), -G(i - 1, j - 1, k + 1), -G(i - 1, j + 1, k - 1), -G(i + 1, j - 1, k - 1),
		-G(i - 1, j + 1, k + 1), -G(i + 1, j - 1, k + 1), -G(i + 1, j + 1, k - 1), -G(i + 1, j + 1, k + 1) };
	return comp_sum(t, 36) / (4 * kPi * tau);
}

// zz: the plane axis is the first; evaluated at (0, j, i)
__device__ double L36_zz(const G3& G, const G3& Dl, int j, int k, double tau)
{
	const int i = 1; j++; k++;
	auto D = [&](int b, int c) { return Dl(0, b, c); };
	double t[36] = {
		+4 * G(i, j, k), +4 * D(j, k),
		-4 * G(i + 1, j, k), -4 * G(i - 1, j, k),
		-2 * G(i, j, k + 1), -2 * D(j, k + 1), -2 * G(i, j, k - 1), -2 * D(j, k - 1),
		-2 * G(i, j + 1, k), -2 * D(j + 1, k), -2 * G(i, j - 1, k), -2 * D(j - 1, k),
		+2 * G(i + 1, j + 1, k), +2 * G(i + 1, j - 1, k), +2 * G(i - 1, j + 1, k), +2 * G(i - 1, j - 1, k),
		+2 * G(i + 1, j, k + 1), +2 * G(i + 1, j, k - 1), +2 * G(i - 1, j, k + 1), +2 * G(i - 1, j, k - 1),
		+G(i, j + 1, k + 1), +D(j + 1, k + 1), +G(i, j + 1, k - 1), +D(j + 1, k - 1),
		+G(i, j - 1, k + 1), +D(j - 1, k + 1), +G(i, j - 1, k - 1), +D(j - 1, k - 1),
		-G(i - 1, j - 1, k - 1), -G(i - 1, j - 1, k + 1), -G(i - 1, j + 1, k - 1), -G(i + 1, j - 1, k - 1),
		-G(i - 1, j + 1, k + 1), -G(i + 1, j - 1, k + 1), -G(i + 1, j + 1, k - 1), -G(i + 1, j + 1, k + 1) };
	return comp_sum(t, 36) / (4 * kPi * tau);
}

// xz, yz: the plane axis is the second; evaluated at (i, 0, k)
__device__ double L36_plane2(const G3& G, const G3& Dl, int i, int k, double tau)
{
	const int j = 1; i++; k++;
	auto D = [&](int a, int c) { return Dl(a, 0, c); };
	double t[36] = {
		+4 * G(i, j, k), +4 * D(i, k),
		-2 * G(i + 1, j, k), -2 * D(i + 1, k), -2 * G(i - 1, j, k), -2 * D(i - 1, k),
		-4 * G(i, j + 1, k), -4 * G(i, j - 1, k),
		-2 * G(i, j, k + 1), -2 * D(i, k + 1), -2 * G(i, j, k - 1), -2 * D(i, k - 1),
		+2 * G(i + 1, j + 1, k), +2 * G(i + 1, j - 1, k), +2 * G(i - 1, j + 1, k), +2 * G(i - 1, j - 1, k),
		+G(i + 1, j, k + 1), +D(i + 1, k + 1), +G(i + 1, j, k - 1), +D(i + 1, k - 1),
		+G(i - 1, j, k + 1), +D(i - 1, k + 1), +G(i - 1, j, k - 1), +D(i - 1, k - 1),
		+2 * G(i, j + 1, k + 1), +2 * G(i, j + 1, k - 1), +2 * G(i, j - 1, k + 1), +2 * G(i, j - 1, k - 1),
		-G(i - 1, j - 1, k - 1), -G(i - 1, j - 1, k + 1), -G(i - 1, j + 1, k - 1), -G(i + 1, j - 1, k - 1),
		-G(i - 1, j + 1, k + 1), -G(i + 1, j - 1, k + 1), -G(i + 1, j + 1, k - 1), -G(i + 1, j + 1, k + 1) };
	return comp_sum(t, 36) / (4 * kPi * tau);
}

struct ZArgs {
	int nx, ny, Nx, Ny, regular;
	int px, py;               // periodic images along x and y (N = n along a periodic axis)
	double sx, sy;            // in-plane shift (k_tensor_shift), normalised
	int lox, loy;             // k_tensor_shift: first offset held by the centred grids, minus one
	double hx, hy, hz, shz, ks, tau, far;
	G3 g[6], del[6];          // f_xx, f_yy, f_zz, g_xy, g_xz, g_yz (+ their extra planes)
	AsymDiag ad[3];
	AsymOff ao[3];
};

// z-shifted 2D tensor (-N convention) into the 6 planes D[q][Ny][Nx] (xx, yy, zz, xy, xz,
// yz): near field by the 27-term (equal cells) or 36-term (irregular) stencils, far field
// (distance >= 40 cells) by the equal-cell series of the destination cell scaled by the
// source thickness for irregular pairs, as the reference does; xy plane mirror images
// with the reference's write order (Boris/DemagTFunc_Shifted.cpp:235-449,
// Boris/DemagTFunc_Irregular.cpp:5-235).
// With periodic images (px or py > 0): every offset i0 (j0) in [0, N) along a periodic axis
// sums the images i = i_img N + i0, i_img = -p..p (x images outer), each at |i| by the
// stencils (grids out to the far-field distance) or the far-field series, the off-diagonal
// terms signed by sign(i) sign(j) (xy), sign(i) (xz), sign(j) (yz) with sign(0) = 0; an open
// axis is mirrored as without images (Boris/DemagTFunc_Shifted_PBC.cpp:317-600,
// Boris/DemagTFunc_Irregular_PBC.cpp:5-235, z-shift branch).
__device__ __forceinline__ int zsign0(int v) { return (v > 0) - (v < 0); }

__device__ __forceinline__ void zshift_term(const ZArgs& a, int i, int j, double vd[3], double vo[3])
{
	const int A = kAsymDist;
	const double hx = a.hx, hy = a.hy, hz = a.hz, sz = a.shz;
	if (i >= A || j >= A || (int)floor(a.ks + 1e-4) >= A || (int)floor(i * i + j * j + a.ks * a.ks + 1e-4) >= A * A) {
		vd[0] = (-asym_diag(a.ad[0], i * hx, j * hy, sz)) * a.far;
		vd[1] = (-asym_diag(a.ad[1], j * hy, i * hx, sz)) * a.far;
		vd[2] = (-asym_diag(a.ad[2], sz, j * hy, i * hx)) * a.far;
		vo[0] = (-asym_off(a.ao[0], i * hx, j * hy, sz)) * a.far;
		vo[1] = (-asym_off(a.ao[1], i * hx, sz, j * hy)) * a.far;
		vo[2] = (-asym_off(a.ao[2], j * hy, sz, i * hx)) * a.far;
	}
	else if (a.regular) {
		vd[0] = -stencil27(a.g[0].v, a.g[0].gx, a.g[0].gy, i, j, 0, hx, hy, hz);
		vd[1] = -stencil27(a.g[1].v, a.g[1].gx, a.g[1].gy, j, i, 0, hy, hx, hz);
		vd[2] = -stencil27(a.g[2].v, a.g[2].gx, a.g[2].gy, 0, j, i, hz, hy, hx);
		vo[0] = -stencil27(a.g[3].v, a.g[3].gx, a.g[3].gy, i, j, 0, hx, hy, hz);
		vo[1] = -stencil27(a.g[4].v, a.g[4].gx, a.g[4].gy, i, 0, j, hx, hz, hy);
		vo[2] = -stencil27(a.g[5].v, a.g[5].gx, a.g[5].gy, j, 0, i, hy, hz, hx);
	}
	else {
		vd[0] = -L36_plane3(a.g[0], a.del[0], i, j, a.tau);
		vd[1] = -L36_plane3(a.g[1], a.del[1], j, i, a.tau);
		vd[2] = -L36_zz(a.g[2], a.del[2], j, i, a.tau);
		vo[0] = -L36_plane3(a.g[3], a.del[3], i, j, a.tau);
		vo[1] = -L36_plane2(a.g[4], a.del[4], i, j, a.tau);
		vo[2] = -L36_plane2(a.g[5], a.del[5], j, i, a.tau);
	}
}

__global__ void k_tensor_zshift_pbc(ZArgs a, double* D)
{
	const size_t NN = (size_t)a.Nx * a.Ny;
	// offsets: [0, n) per axis, i.e. [0, N) along a periodic axis and [0, N/2) along an open one
	for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < (size_t)a.nx * a.ny; q += (size_t)gridDim.x * blockDim.x) {
		const int i0 = (int)(q % a.nx), j0 = (int)(q / a.nx);
		double vd[3] = { 0.0, 0.0, 0.0 }, vo[3] = { 0.0, 0.0, 0.0 };
		for (int ii = -a.px; ii <= a.px; ii++)
			for (int jj = -a.py; jj <= a.py; jj++) {
				const int is = ii * a.Nx + i0, js = jj * a.Ny + j0;
				const double si = zsign0(is), sj = zsign0(js);
				double td[3], to[3];
				zshift_term(a, abs(is), abs(js), td, to);
				vd[0] += td[0]; vd[1] += td[1]; vd[2] += td[2];
				vo[0] += to[0] * (si * sj); vo[1] += to[1] * si; vo[2] += to[2] * sj;
			}
		auto put = [&](size_t idx, double ox, double oy, double oz) {
			D[idx] = vd[0]; D[NN + idx] = vd[1]; D[2 * NN + idx] = vd[2];
			D[3 * NN + idx] = vo[0] * ox; D[4 * NN + idx] = vo[1] * oy; D[5 * NN + idx] = vo[2] * oz;
		};
		const size_t mx = (size_t)((a.Nx - i0) % a.Nx), my = (size_t)((a.Ny - j0) % a.Ny);
		put((size_t)i0 + (size_t)j0 * a.Nx, 1, 1, 1);
		if (!a.px) {
			if (i0) put(mx + (size_t)j0 * a.Nx, -1, -1, 1);
			if (!a.py && i0 && j0) put(mx + my * a.Nx, 1, -1, -1);
		}
		if (!a.py && j0) put((size_t)i0 + my * a.Nx, -1, 1, -1);
	}
}

// Tensor shifted in-plane (sx or sy nonzero) and along z: every offset i in (-nx, nx), j in
// (-ny, ny) at position (i hx + sx, j hy + sy, shz), no symmetry, stored at (i mod Nx, j mod Ny)
// (the general branch of Boris/DemagTFunc_Shifted.cpp:235-343 and
// Boris/DemagTFunc_Irregular.cpp:5-235): near field by the same stencils on grids centred on
// the offsets the near field reads, far field by the series at the shifted position.
__global__ void k_tensor_shift(ZArgs a, double* D)
{
	const size_t NN = (size_t)a.Nx * a.Ny;
	const int A = kAsymDist, W = 2 * a.nx - 1, Hh = 2 * a.ny - 1;
	const double hx = a.hx, hy = a.hy, hz = a.hz, sz = a.shz;
	for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < (size_t)W * Hh; q += (size_t)gridDim.x * blockDim.x) {
		const int i = (int)(q % W) - (a.nx - 1), j = (int)(q / W) - (a.ny - 1);
		const double is = fabs(i + a.sx / hx), js = fabs(j + a.sy / hy), ks = a.ks;
		const double x = i * hx + a.sx, y = j * hy + a.sy;
		double vd[3], vo[3];
		if ((int)floor(is + 1e-4) >= A || (int)floor(js + 1e-4) >= A || (int)floor(ks + 1e-4) >= A || (int)floor(is * is + js * js + ks * ks + 1e-4) >= A * A) {
			vd[0] = (-asym_diag(a.ad[0], x, y, sz)) * a.far;
			vd[1] = (-asym_diag(a.ad[1], y, x, sz)) * a.far;
			vd[2] = (-asym_diag(a.ad[2], sz, y, x)) * a.far;
			vo[0] = (-asym_off(a.ao[0], x, y, sz)) * a.far;
			vo[1] = (-asym_off(a.ao[1], x, sz, y)) * a.far;
			vo[2] = (-asym_off(a.ao[2], y, sz, x)) * a.far;
		}
		else {
			const int qi = i - a.lox - 1, qj = j - a.loy - 1;
			if (a.regular) {
				vd[0] = -stencil27(a.g[0].v, a.g[0].gx, a.g[0].gy, qi, qj, 0, hx, hy, hz);
				vd[1] = -stencil27(a.g[1].v, a.g[1].gx, a.g[1].gy, qj, qi, 0, hy, hx, hz);
				vd[2] = -stencil27(a.g[2].v, a.g[2].gx, a.g[2].gy, 0, qj, qi, hz, hy, hx);
				vo[0] = -stencil27(a.g[3].v, a.g[3].gx, a.g[3].gy, qi, qj, 0, hx, hy, hz);
				vo[1] = -stencil27(a.g[4].v, a.g[4].gx, a.g[4].gy, qi, 0, qj, hx, hz, hy);
				vo[2] = -stencil27(a.g[5].v, a.g[5].gx, a.g[5].gy, qj, 0, qi, hy, hz, hx);
			}
			else {
				vd[0] = -L36_plane3(a.g[0], a.del[0], qi, qj, a.tau);
				vd[1] = -L36_plane3(a.g[1], a.del[1], qj, qi, a.tau);
				vd[2] = -L36_zz(a.g[2], a.del[2], qj, qi, a.tau);
				vo[0] = -L36_plane3(a.g[3], a.del[3], qi, qj, a.tau);
				vo[1] = -L36_plane2(a.g[4], a.del[4], qi, qj, a.tau);
				vo[2] = -L36_plane2(a.g[5], a.del[5], qj, qi, a.tau);
			}
		}
		const size_t idx = (size_t)((i + a.Nx) % a.Nx) + (size_t)((j + a.Ny) % a.Ny) * a.Nx;
		D[idx] = vd[0]; D[NN + idx] = vd[1]; D[2 * NN + idx] = vd[2];
		D[3 * NN + idx] = vo[0]; D[4 * NN + idx] = vo[1]; D[5 * NN + idx] = vo[2];
	}
}

// complex kernel planes from the y FFTs of the x-transformed tensor's real (A) and imaginary
// (B) parts: Re = Re A - Im B, Im = Im A + Re B
__global__ void k_cplx_combine(const double* Are, const double* Aim, const double* Bre, const double* Bim, float* re, float* im, size_t n)
{
	for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x) {
		re[q] = (float)(Are[q] - Bim[q]);
		im[q] = (float)(Aim[q] + Bre[q]);
	}
}

__global__ void k_tensor_zshift(ZArgs a, double* D)
{
	const size_t NN = (size_t)a.Nx * a.Ny;
	const int A = kAsymDist;
	const int mir[4][2] = { {1,1},{-1,1},{1,-1},{-1,-1} };
	for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < (size_t)a.nx * a.ny; q += (size_t)gridDim.x * blockDim.x) {
		const int i = (int)(q % a.nx), j = (int)(q / a.nx);
		const double hx = a.hx, hy = a.hy, hz = a.hz, sz = a.shz;
		double vd[3], vo[3];
		if (i >= A || j >= A || (int)floor(a.ks + 1e-4) >= A || (int)floor(i * i + j * j + a.ks * a.ks + 1e-4) >= A * A) {
			vd[0] = (-asym_diag(a.ad[0], i * hx, j * hy, sz)) * a.far;
			vd[1] = (-asym_diag(a.ad[1], j * hy, i * hx, sz)) * a.far;
			vd[2] = (-asym_diag(a.ad[2], sz, j * hy, i * hx)) * a.far;
			vo[0] = (-asym_off(a.ao[0], i * hx, j * hy, sz)) * a.far;
			vo[1] = (-asym_off(a.ao[1], i * hx, sz, j * hy)) * a.far;
			vo[2] = (-asym_off(a.ao[2], j * hy, sz, i * hx)) * a.far;
		}
		else if (a.regular) {
			vd[0] = -stencil27(a.g[0].v, a.g[0].gx, a.g[0].gy, i, j, 0, hx, hy, hz);
			vd[1] = -stencil27(a.g[1].v, a.g[1].gx, a.g[1].gy, j, i, 0, hy, hx, hz);
			vd[2] = -stencil27(a.g[2].v, a.g[2].gx, a.g[2].gy, 0, j, i, hz, hy, hx);
			vo[0] = -stencil27(a.g[3].v, a.g[3].gx, a.g[3].gy, i, j, 0, hx, hy, hz);
			vo[1] = -stencil27(a.g[4].v, a.g[4].gx, a.g[4].gy, i, 0, j, hx, hz, hy);
			vo[2] = -stencil27(a.g[5].v, a.g[5].gx, a.g[5].gy, j, 0, i, hy, hz, hx);
		}
		else {
			vd[0] = -L36_plane3(a.g[0], a.del[0], i, j, a.tau);
			vd[1] = -L36_plane3(a.g[1], a.del[1], j, i, a.tau);
			vd[2] = -L36_zz(a.g[2], a.del[2], j, i, a.tau);
			vo[0] = -L36_plane3(a.g[3], a.del[3], i, j, a.tau);
			vo[1] = -L36_plane2(a.g[4], a.del[4], i, j, a.tau);
			vo[2] = -L36_plane2(a.g[5], a.del[5], j, i, a.tau);
		}
		for (int m = 0; m < 4; m++) {
			bool later = false;
			for (int p = m + 1; p < 4; p++)
				if ((mir[p][0] == mir[m][0] || i == 0) && (mir[p][1] == mir[m][1] || j == 0)) later = true;
			if (later) continue;
			const int si = mir[m][0], sj = mir[m][1];
			const size_t idx = (size_t)((si * i + a.Nx) % a.Nx) + (size_t)((sj * j + a.Ny) % a.Ny) * a.Nx;
			D[idx] = vd[0]; D[NN + idx] = vd[1]; D[2 * NN + idx] = vd[2];
			D[3 * NN + idx] = vo[0] * (si * sj);
			D[4 * NN + idx] = vo[1] * si;
			D[5 * NN + idx] = vo[2] * sj;
		}
	}
}

// Kernel sums of all destinations at one (kx, ky) point:
//   out_d = sum_s [Kxx Fx + Kxy Fy + i Kxz Fz,  Kxy Fx + Kyy Fy + i Kyz Fz,  i Kxz Fx + i Kyz Fy + Kzz Fz]_ds
// with the kernels of (kx, |ky| folded to <= Ny/2) and Kxy, Kyz negated above Ny/2 (odd in
// y); the own layer first, then the others in index order (Boris/DemagKernelCollection_Mult.cpp:457-500).
// Self kernels have Kxz = Kyz = 0.  In place: all layers' inputs are read before any output.
// tr: the transposed operator (N^T)_sd = N_ds^T (the adjoint of the convolution): output s
// sums the pair kernels K_ds over inputs d with the imaginary xz, yz parts negated (the
// real-space kernel N_ds(-r)^T).
// K holds one kernel of 6 planes per pair (d, s) at d * L + s.
template <int L>
__global__ void __launch_bounds__(256) k_ml_mult(float2* __restrict__ S, const float* __restrict__ K, int NC, int Ny, int Sp, int tr,
	const MlBatch* __restrict__ bt)
{
	if (bt) { const MlBatch& b = bt[blockIdx.y]; S = b.S; K = b.K; NC = b.NC; Ny = b.Ny; Sp = b.Sp; }   // batched: one stack per grid row
	const int NYH = Ny / 2 + 1;
	const size_t kq = (size_t)NC * NYH;
	const int npts = NC * Ny;
	for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < npts; q += gridDim.x * blockDim.x) {
		const int ky = q % Ny, kx = q / Ny;
		const bool up = ky > Ny / 2;
		const int jj = up ? Ny - ky : ky;
		const float sy = up ? -1.f : 1.f;
		float2 F[L][3];
#pragma unroll
		for (int s = 0; s < L; s++)
#pragma unroll
			for (int c = 0; c < 3; c++) F[s][c] = S[((size_t)(c * L + s) * NC + kx) * Sp + ky];
#pragma unroll
		for (int d = 0; d < L; d++) {
			float2 o0 = make_float2(0.f, 0.f), o1 = o0, o2 = o0;
#pragma unroll
			for (int r = 0; r < L; r++) {
				const int s = r == 0 ? d : (r <= d ? r - 1 : r);
				const float* Kp = K + (size_t)(tr ? s * L + d : d * L + s) * 6 * kq + (size_t)kx * NYH + jj;
				const float tz = tr ? -1.f : 1.f;
				const float kxx = Kp[0], kyy = Kp[kq], kzz = Kp[2 * kq];
				const float kxy = sy * Kp[3 * kq], kxz = tz * Kp[4 * kq], kyz = tz * sy * Kp[5 * kq];
				const float2 fx = F[s][0], fy = F[s][1], fz = F[s][2];
				// i k z = k * (-z.y, z.x)
				o0.x += kxx * fx.x + kxy * fy.x - kxz * fz.y;  o0.y += kxx * fx.y + kxy * fy.y + kxz * fz.x;
				o1.x += kxy * fx.x + kyy * fy.x - kyz * fz.y;  o1.y += kxy * fx.y + kyy * fy.y + kyz * fz.x;
				o2.x += -kxz * fx.y - kyz * fy.y + kzz * fz.x; o2.y += kxz * fx.x + kyz * fy.x + kzz * fz.y;
			}
			S[((size_t)(0 * L + d) * NC + kx) * Sp + ky] = o0;
			S[((size_t)(1 * L + d) * NC + kx) * Sp + ky] = o1;
			S[((size_t)(2 * L + d) * NC + kx) * Sp + ky] = o2;
		}
	}
}

// More than kRegLayers layers: the same sums (own layer first, then the others in index order)
// for DG destinations at a time, the sources read back from S for every group, the results
// written to S2 (out of place, since later groups still read the inputs).  Kernels by slot:
// kidx[d * L + s] is the slot of pair (d, s) (pairs of equal thicknesses and z shift share
// one slot of 6 planes).
constexpr int kDG = 8;
__global__ void __launch_bounds__(256) k_ml_mult_gen(const float2* __restrict__ S, float2* __restrict__ S2, const float* __restrict__ K,
	const int* __restrict__ kidx, int L, int NC, int Ny, int Sp, int tr)
{
	const int NYH = Ny / 2 + 1;
	const size_t kq = (size_t)NC * NYH;
	const int npts = NC * Ny;
	const float tz = tr ? -1.f : 1.f;
	for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < npts; q += gridDim.x * blockDim.x) {
		const int ky = q % Ny, kx = q / Ny;
		const bool up = ky > Ny / 2;
		const int jj = up ? Ny - ky : ky;
		const float sy = up ? -1.f : 1.f;
		auto F = [&](int s, int c) { return S[((size_t)(c * L + s) * NC + kx) * Sp + ky]; };
		for (int d0 = 0; d0 < L; d0 += kDG) {
			float2 o[kDG][3];
#pragma unroll
			for (int g = 0; g < kDG; g++) for (int c = 0; c < 3; c++) o[g][c] = make_float2(0.f, 0.f);
			auto acc = [&](int g, int d, int s) {
				const float* Kp = K + (size_t)kidx[tr ? s * L + d : d * L + s] * 6 * kq + (size_t)kx * NYH + jj;
				const float kxx = Kp[0], kyy = Kp[kq], kzz = Kp[2 * kq];
				const float kxy = sy * Kp[3 * kq], kxz = tz * Kp[4 * kq], kyz = tz * sy * Kp[5 * kq];
				const float2 fx = F(s, 0), fy = F(s, 1), fz = F(s, 2);
				o[g][0].x += kxx * fx.x + kxy * fy.x - kxz * fz.y;  o[g][0].y += kxx * fx.y + kxy * fy.y + kxz * fz.x;
				o[g][1].x += kxy * fx.x + kyy * fy.x - kyz * fz.y;  o[g][1].y += kxy * fx.y + kyy * fy.y + kyz * fz.x;
				o[g][2].x += -kxz * fx.y - kyz * fy.y + kzz * fz.x; o[g][2].y += kxz * fx.x + kyz * fy.x + kzz * fz.y;
			};
#pragma unroll
			for (int g = 0; g < kDG; g++) if (d0 + g < L) acc(g, d0 + g, d0 + g);
			for (int s = 0; s < L; s++) {
#pragma unroll
				for (int g = 0; g < kDG; g++) if (d0 + g < L && s != d0 + g) acc(g, d0 + g, s);
			}
#pragma unroll
			for (int g = 0; g < kDG; g++)
				if (d0 + g < L)
					for (int c = 0; c < 3; c++) S2[((size_t)(c * L + d0 + g) * NC + kx) * Sp + ky] = o[g][c];
		}
	}
}

// The kernel product; the result is in S (L <= kRegLayers, in place) or S2.
// Layers offset in-plane against each other (their pairs have complex kernels): the same sums
// (own layer first, then the others in index order), a pair with kcoff[p] >= 0 through its
// complex kernel (12 planes Re, Im of xx, yy, zz, xy, xz, yz, NC x Ny each, ky fastest, at
// Kc + kcoff[p]; Boris/DemagKernelCollectionCUDA_Mult.cu's Complex_Full product), the others
// through their real kernels (K, per pair or by slot kidx) as in k_ml_mult; out of place into S2.
__global__ void __launch_bounds__(256) k_ml_mult_any(const float2* __restrict__ S, float2* __restrict__ S2, const float* __restrict__ K,
	const int* __restrict__ kidx, const float* __restrict__ Kc, const long long* __restrict__ kcoff, int L, int NC, int Ny, int Sp)
{
	const int NYH = Ny / 2 + 1;
	const size_t kq = (size_t)NC * NYH, kc = (size_t)NC * Ny;
	const int npts = NC * Ny;
	auto cm = [](float2 k, float2 f) { return make_float2(k.x * f.x - k.y * f.y, k.x * f.y + k.y * f.x); };
	for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < npts; q += gridDim.x * blockDim.x) {
		const int ky = q % Ny, kx = q / Ny;
		const bool up = ky > Ny / 2;
		const int jj = up ? Ny - ky : ky;
		const float sy = up ? -1.f : 1.f;
		for (int d = 0; d < L; d++) {
			float2 o0 = make_float2(0.f, 0.f), o1 = o0, o2 = o0;
			for (int r = 0; r < L; r++) {
				const int s = r == 0 ? d : (r <= d ? r - 1 : r);
				const int p = d * L + s;
				const float2 fx = S[((size_t)(0 * L + s) * NC + kx) * Sp + ky], fy = S[((size_t)(1 * L + s) * NC + kx) * Sp + ky];
				const float2 fz = S[((size_t)(2 * L + s) * NC + kx) * Sp + ky];
				if (kcoff[p] >= 0) {
					const float* Kp = Kc + kcoff[p] + (size_t)kx * Ny + ky;
					auto k = [&](int c) { return make_float2(Kp[(2 * c) * kc], Kp[(2 * c + 1) * kc]); };
					const float2 kxx = k(0), kyy = k(1), kzz = k(2), kxy = k(3), kxz = k(4), kyz = k(5);
					const float2 a0 = cm(kxx, fx), a1 = cm(kxy, fy), a2 = cm(kxz, fz);
					const float2 b0 = cm(kxy, fx), b1 = cm(kyy, fy), b2 = cm(kyz, fz);
					const float2 c0 = cm(kxz, fx), c1 = cm(kyz, fy), c2 = cm(kzz, fz);
					o0.x += a0.x + a1.x + a2.x; o0.y += a0.y + a1.y + a2.y;
					o1.x += b0.x + b1.x + b2.x; o1.y += b0.y + b1.y + b2.y;
					o2.x += c0.x + c1.x + c2.x; o2.y += c0.y + c1.y + c2.y;
					continue;
				}
				const float* Kp = K + (size_t)(kidx ? kidx[p] : p) * 6 * kq + (size_t)kx * NYH + jj;
				const float kxx = Kp[0], kyy = Kp[kq], kzz = Kp[2 * kq];
				const float kxy = sy * Kp[3 * kq], kxz = Kp[4 * kq], kyz = sy * Kp[5 * kq];
				o0.x += kxx * fx.x + kxy * fy.x - kxz * fz.y;  o0.y += kxx * fx.y + kxy * fy.y + kxz * fz.x;
				o1.x += kxy * fx.x + kyy * fy.x - kyz * fz.y;  o1.y += kxy * fx.y + kyy * fy.y + kyz * fz.x;
				o2.x += -kxz * fx.y - kyz * fy.y + kzz * fz.x; o2.y += kxz * fx.x + kyz * fy.x + kzz * fz.y;
			}
			S2[((size_t)(0 * L + d) * NC + kx) * Sp + ky] = o0;
			S2[((size_t)(1 * L + d) * NC + kx) * Sp + ky] = o1;
			S2[((size_t)(2 * L + d) * NC + kx) * Sp + ky] = o2;
		}
	}
}

int launch_ml_mult(int L, float2* S, float2* S2, const float* K, const int* kidx, int NC, int Ny, int Sp, int ncu, hipStream_t st, int tr = 0,
	const float* Kc = nullptr, const long long* kcoff = nullptr)
{
	const int npts = NC * Ny;
	dim3 g((unsigned)std::max(1, std::min((npts + 255) / 256, ncu * 8))), b(256);
	if (Kc) {
		if (!S2 || !kcoff) return fail(BM_ERR_STATE, "multilayer demag: no second spectra array");
		if (tr) return fail(BM_ERR_STATE, "multilayer demag: no transposed product for layers offset in-plane");
		hipLaunchKernelGGL(k_ml_mult_any, g, b, 0, st, S, S2, K, kidx, Kc, kcoff, L, NC, Ny, Sp);
		return BM_OK;
	}
	if (L > kRegLayers) {
		if (!S2) return fail(BM_ERR_STATE, "multilayer demag: no second spectra array");
		hipLaunchKernelGGL(k_ml_mult_gen, g, b, 0, st, S, S2, K, kidx, L, NC, Ny, Sp, tr);
		return BM_OK;
	}
	switch (L) {
#define C_(V_) case V_: hipLaunchKernelGGL((k_ml_mult<V_>), g, b, 0, st, S, K, NC, Ny, Sp, tr, (const MlBatch*)nullptr); return BM_OK;
	C_(1) C_(2) C_(3) C_(4) C_(5) C_(6) C_(7) C_(8) C_(9) C_(10) C_(11) C_(12) C_(13) C_(14) C_(15) C_(16)
#undef C_
	}
	return fail(BM_ERR_SIZE, "multilayer demag: layer count");
}

} // namespace

struct bm_mlayer_s {
	int L, nx, ny, Nx, Ny, NC, NYH, Sp, ntl;
	int xs = 0;                  // x-pass size N.x / 2 (nx, or nx / 2 along a periodic x)
	int pbc[2] = { 0, 0 };       // periodic images along x and y (bm_mlayer_create_pbc)
	bool periodic() const { return pbc[0] || pbc[1]; }
	// layers whose grids are offset in-plane by fractions of a cell (rx, ry per layer, m; empty
	// otherwise): their pairs couple through complex kernels, 12 planes of NC x Ny floats each
	// at Kc + hkcoff[d * L + s] (-1: the real kernel in K)
	std::vector<double> rx, ry;
	std::vector<long long> hkcoff;
	float* Kc = nullptr;
	long long* kcoff = nullptr;
	double hx, hy;
	std::vector<double> hz, z0;
	float2* Sx = nullptr;    // x-side spectra, row-tiled (bm_xpass.h): 3 * L * ntl * NC * kLT complex
	float2* S = nullptr;     // full spectra S[c][layer][kx][ky]: 3 * L * NC * Ny complex
	float2* S2 = nullptr;    // L > kRegLayers: the product's output (same layout), else null
	// kernels, 6 planes of NC x (Ny/2+1) each, ky fastest: one per pair (d * L + s) for
	// L <= kRegLayers; above, one per slot (pairs with equal thicknesses and z shift share one)
	float* K = nullptr;
	int nslot = 0;
	std::vector<int> hkidx;  // slot of pair (d, s) at d * L + s
	int* kidx = nullptr;     // its device copy (L > kRegLayers)
	bool per_pair() const { return L <= kRegLayers; }
	size_t kbase(int d, int s) const { return (size_t)(per_pair() ? d * L + s : hkidx[(size_t)d * L + s]); }
	long long kcount = 0, bytes = 0;
	int ncu = 256;
	double* ebuf = nullptr;  // energy partials of the per-layer inverse (bm_mlayer_inverse)
	float2* Sout() const { return S2 ? S2 : S; }
	// layers that are meshes of their own (bm_mlayer_create_meshes) on the common grid above:
	// empty when every layer is the common grid itself
	std::vector<long long> moff;   // first cell of each layer mesh in the concatenation (L + 1)
	void* tabs = nullptr;          // device transfer tables (in: L * plane rows, out: moff[L] rows)
	Csr tin{}, tout{};
	float* stage = nullptr;        // (L + 1) common planes, AoS: the layers' M (then H), one H plane
	double* epart = nullptr;       // energy partials of the per-layer transfer out
	bool mapped() const { return !moff.empty(); }
};

namespace {

// z-shifted kernels of pair (d, s) into Kp (6 planes of NC*NYH floats).
int zshift_kernels(bm_mlayer_s* m, int d, int s, double hmax, double* D, double* T1, double* grids[12], float* Kp, hipStream_t st)
{
	const int nx = m->nx, ny = m->ny, Nx = m->Nx, Ny = m->Ny, NC = m->NC, NYH = m->NYH;
	const double sc[3] = { m->hx / hmax, m->hy / hmax, m->hz[s] / hmax }, dc[3] = { m->hx / hmax, m->hy / hmax, m->hz[d] / hmax };
	const double shz = (m->z0[d] - m->z0[s]) / hmax;
	bool regular = true;
	for (int c = 0; c < 3; c++) if (nzero(sc[c] * 1e-9 - dc[c] * 1e-9)) regular = false;
	// the regular path runs on the source cell (Boris/DemagTFunc_Irregular.cpp:11)
	const double hx = regular ? sc[0] : dc[0], hy = regular ? sc[1] : dc[1], hz = regular ? sc[2] : dc[2];
	// near-field grids: out to the far-field distance along a periodic axis (images reach it)
	const int dx = m->pbc[0] ? kAsymDist : std::min(nx, kAsymDist), dy = m->pbc[1] ? kAsymDist : std::min(ny, kAsymDist);
	struct Spec { int which, gx, gy, gz; Axis X, Y, Z; };
	Spec sp[12];
	if (regular) {
		sp[0] = { 0, dx, dy, 1, lin(hx, 0), lin(hy, 0), lin(hz, shz) };
		sp[1] = { 0, dy, dx, 1, lin(hy, 0), lin(hx, 0), lin(hz, shz) };
		sp[2] = { 0, 1, dy, dx, lin(hz, shz), lin(hy, 0), lin(hx, 0) };
		sp[3] = { 1, dx, dy, 1, lin(hx, 0), lin(hy, 0), lin(hz, shz) };
		sp[4] = { 1, dx, 1, dy, lin(hx, 0), lin(hz, shz), lin(hy, 0) };
		sp[5] = { 1, dy, 1, dx, lin(hy, 0), lin(hz, shz), lin(hx, 0) };
	}
	else {
		// planes -1, 0, +1 at -s.z, 0, +d.z (xx, yy) and -d.z, 0, +s.z (zz and g), plus the
		// extra plane at -del (xx, yy) or +del (zz and g), all offset by the shift
		// (Boris/DemagTFunc_fg.cpp:507-620)
		const double del = sc[2] - dc[2];
		const Axis p_sd = tab(-sc[2] + shz, shz, dc[2] + shz), p_ds = tab(-dc[2] + shz, shz, sc[2] + shz);
		const Axis c_m = tab(-del + shz, 0, 0), c_p = tab(del + shz, 0, 0);
		sp[0] = { 0, dx, dy, 1, lin(hx, 0), lin(hy, 0), p_sd };
		sp[1] = { 0, dy, dx, 1, lin(hy, 0), lin(hx, 0), p_sd };
		sp[2] = { 0, 1, dy, dx, p_ds, lin(hy, 0), lin(hx, 0) };
		sp[3] = { 1, dx, dy, 1, lin(hx, 0), lin(hy, 0), p_ds };
		sp[4] = { 1, dx, 1, dy, lin(hx, 0), p_ds, lin(hy, 0) };
		sp[5] = { 1, dy, 1, dx, lin(hy, 0), p_ds, lin(hx, 0) };
		sp[6] = { 0, dx, dy, -1, lin(hx, 0), lin(hy, 0), c_m };
		sp[7] = { 0, dy, dx, -1, lin(hy, 0), lin(hx, 0), c_m };
		sp[8] = { 0, -1, dy, dx, c_p, lin(hy, 0), lin(hx, 0) };
		sp[9] = { 1, dx, dy, -1, lin(hx, 0), lin(hy, 0), c_p };
		sp[10] = { 1, dx, -1, dy, lin(hx, 0), c_p, lin(hy, 0) };
		sp[11] = { 1, dy, -1, dx, lin(hy, 0), c_p, lin(hx, 0) };
	}
	ZArgs a;
	std::memset(&a, 0, sizeof(a));
	a.nx = nx; a.ny = ny; a.Nx = Nx; a.Ny = Ny; a.regular = regular ? 1 : 0;
	a.px = m->pbc[0]; a.py = m->pbc[1];
	a.hx = hx; a.hy = hy; a.hz = hz; a.shz = shz; a.ks = std::fabs(shz / hz); a.tau = hx * hy * hz;
	a.far = regular ? 1.0 : sc[2];
	const int ng = regular ? 6 : 12;
	for (int q = 0; q < ng; q++) {
		const size_t cnt = (size_t)(sp[q].gx + 2) * (sp[q].gy + 2) * (sp[q].gz + 2);
		hipLaunchKernelGGL(k_fill_grid_at, dim3((unsigned)std::min<size_t>((cnt + 255) / 256, 4096)), dim3(256), 0, st,
			grids[q], sp[q].which, sp[q].gx, sp[q].gy, sp[q].gz, sp[q].X, sp[q].Y, sp[q].Z);
		G3& G = q < 6 ? a.g[q] : a.del[q - 6];
		G.v = grids[q]; G.gx = sp[q].gx + 2; G.gy = sp[q].gy + 2;
	}
	asym_diag_init(a.ad[0], hx, hy, hz); asym_diag_init(a.ad[1], hy, hx, hz); asym_diag_init(a.ad[2], hz, hy, hx);
	asym_off_init(a.ao[0], hx, hy, hz); asym_off_init(a.ao[1], hx, hz, hy); asym_off_init(a.ao[2], hy, hz, hx);
	const size_t NN = (size_t)Nx * Ny;
	BM_HIP(hipMemsetAsync(D, 0, NN * 6 * sizeof(double), st));
	const dim3 tg((unsigned)std::min<size_t>(((size_t)nx * ny + 127) / 128, 65535));
	if (a.px || a.py) hipLaunchKernelGGL(k_tensor_zshift_pbc, tg, dim3(128), 0, st, a, D);
	else hipLaunchKernelGGL(k_tensor_zshift, tg, dim3(128), 0, st, a, D);
	// x then y real-input FFTs with the part each component keeps
	// (Boris/DemagKernelCollection_Calc.cpp:288-420): diagonal Re, Re; xy Im, -Im; xz Im, Re; yz Re, Im
	static const int px[6] = { 0, 0, 0, 1, 1, 0 }, py[6] = { 0, 0, 0, 1, 0, 1 };
	static const double sg[6] = { 1, 1, 1, -1, 1, 1 };
	for (int q = 0; q < 6; q++) {
		int rc = setup_line_fft(Nx, D + q * NN, T1, Ny, Ny, Nx, 0, 1, NC, 0, 1, NC, px[q], 1.0, st);
		if (rc == BM_OK) rc = setup_line_fft(Ny, T1, Kp + (size_t)q * NC * NYH, NC, NC, 1, 0, NC, NYH, 0, 1, NYH, py[q], sg[q], st);
		if (rc) return rc;
	}
	return BM_OK;
}

// Complex kernels of pair (d, s) whose layers are offset in-plane (rx, ry differ): the tensor
// shifted by the offset difference (k_tensor_shift), x real-input FFT keeping both parts, then
// the complex y FFT as the real-input FFTs of those parts, into 12 planes of NC x Ny floats at
// Kp (Boris/DemagKernelCollectionCUDA_Calc.cpp:571-745).  T holds 6 NC x Ny doubles.
int shift_kernels(bm_mlayer_s* m, int d, int s, double hmax, double* D, double* T, double* grids[12], float* Kp, hipStream_t st)
{
	const int nx = m->nx, ny = m->ny, Nx = m->Nx, Ny = m->Ny, NC = m->NC;
	const double sc[3] = { m->hx / hmax, m->hy / hmax, m->hz[s] / hmax }, dc[3] = { m->hx / hmax, m->hy / hmax, m->hz[d] / hmax };
	const double shx = (m->rx[d] - m->rx[s]) / hmax, shy = (m->ry[d] - m->ry[s]) / hmax, shz = (m->z0[d] - m->z0[s]) / hmax;
	bool regular = true;
	for (int c = 0; c < 3; c++) if (nzero(sc[c] * 1e-9 - dc[c] * 1e-9)) regular = false;
	const double hx = regular ? sc[0] : dc[0], hy = regular ? sc[1] : dc[1], hz = regular ? sc[2] : dc[2];
	// grids over the offsets the near field reads: |i + sh / h| below the far-field distance, +-1
	auto lo_of = [](int n, double f) { return std::max(-n, (int)std::floor(-kAsymDist - f) - 2); };
	auto hi_of = [](int n, double f) { return std::min(n, (int)std::ceil(kAsymDist - f) + 2); };
	const int lox = lo_of(nx, shx / hx), loy = lo_of(ny, shy / hy);
	const int dx = hi_of(nx, shx / hx) - lox - 1, dy = hi_of(ny, shy / hy) - loy - 1;
	const Axis X = cen(hx, shx, lox), Y = cen(hy, shy, loy);
	struct Spec { int which, gx, gy, gz; Axis X, Y, Z; };
	Spec sp[12];
	if (regular) {
		sp[0] = { 0, dx, dy, 1, X, Y, lin(hz, shz) };
		sp[1] = { 0, dy, dx, 1, Y, X, lin(hz, shz) };
		sp[2] = { 0, 1, dy, dx, lin(hz, shz), Y, X };
		sp[3] = { 1, dx, dy, 1, X, Y, lin(hz, shz) };
		sp[4] = { 1, dx, 1, dy, X, lin(hz, shz), Y };
		sp[5] = { 1, dy, 1, dx, Y, lin(hz, shz), X };
	}
	else {
		const double del = sc[2] - dc[2];
		const Axis p_sd = tab(-sc[2] + shz, shz, dc[2] + shz), p_ds = tab(-dc[2] + shz, shz, sc[2] + shz);
		const Axis c_m = tab(-del + shz, 0, 0), c_p = tab(del + shz, 0, 0);
		sp[0] = { 0, dx, dy, 1, X, Y, p_sd };
		sp[1] = { 0, dy, dx, 1, Y, X, p_sd };
		sp[2] = { 0, 1, dy, dx, p_ds, Y, X };
		sp[3] = { 1, dx, dy, 1, X, Y, p_ds };
		sp[4] = { 1, dx, 1, dy, X, p_ds, Y };
		sp[5] = { 1, dy, 1, dx, Y, p_ds, X };
		sp[6] = { 0, dx, dy, -1, X, Y, c_m };
		sp[7] = { 0, dy, dx, -1, Y, X, c_m };
		sp[8] = { 0, -1, dy, dx, c_p, Y, X };
		sp[9] = { 1, dx, dy, -1, X, Y, c_p };
		sp[10] = { 1, dx, -1, dy, X, c_p, Y };
		sp[11] = { 1, dy, -1, dx, Y, c_p, X };
	}
	ZArgs a;
	std::memset(&a, 0, sizeof(a));
	a.nx = nx; a.ny = ny; a.Nx = Nx; a.Ny = Ny; a.regular = regular ? 1 : 0;
	a.hx = hx; a.hy = hy; a.hz = hz; a.shz = shz; a.ks = std::fabs(shz / hz); a.tau = hx * hy * hz;
	a.far = regular ? 1.0 : sc[2];
	a.sx = shx; a.sy = shy; a.lox = lox; a.loy = loy;
	const int ng = regular ? 6 : 12;
	for (int q = 0; q < ng; q++) {
		const size_t cnt = (size_t)(sp[q].gx + 2) * (sp[q].gy + 2) * (sp[q].gz + 2);
		hipLaunchKernelGGL(k_fill_grid_at, dim3((unsigned)std::min<size_t>((cnt + 255) / 256, 4096)), dim3(256), 0, st,
			grids[q], sp[q].which, sp[q].gx, sp[q].gy, sp[q].gz, sp[q].X, sp[q].Y, sp[q].Z);
		G3& G = q < 6 ? a.g[q] : a.del[q - 6];
		G.v = grids[q]; G.gx = sp[q].gx + 2; G.gy = sp[q].gy + 2;
	}
	asym_diag_init(a.ad[0], hx, hy, hz); asym_diag_init(a.ad[1], hy, hx, hz); asym_diag_init(a.ad[2], hz, hy, hx);
	asym_off_init(a.ao[0], hx, hy, hz); asym_off_init(a.ao[1], hx, hz, hy); asym_off_init(a.ao[2], hy, hz, hx);
	const size_t NN = (size_t)Nx * Ny, nk = (size_t)NC * Ny;
	BM_HIP(hipMemsetAsync(D, 0, NN * 6 * sizeof(double), st));
	const size_t work = (size_t)(2 * nx - 1) * (2 * ny - 1);
	hipLaunchKernelGGL(k_tensor_shift, dim3((unsigned)std::min<size_t>((work + 127) / 128, 65535)), dim3(128), 0, st, a, D);
	double* Tre = T; double* Tim = T + nk; double* Are = T + 2 * nk; double* Aim = T + 3 * nk; double* Bre = T + 4 * nk; double* Bim = T + 5 * nk;
	for (int q = 0; q < 6; q++) {
		// x: both parts of the real-input FFT (kx in [0, Nx/2], layout kx + j NC)
		int rc = setup_line_fft(Nx, D + q * NN, Tre, Ny, Ny, Nx, 0, 1, NC, 0, 1, NC, 0, 1.0, st);
		if (rc == BM_OK) rc = setup_line_fft(Nx, D + q * NN, Tim, Ny, Ny, Nx, 0, 1, NC, 0, 1, NC, 1, 1.0, st);
		// y: every ky (layout kx Ny + ky) of the two real parts
		if (rc == BM_OK) rc = setup_line_fft(Ny, Tre, Are, NC, NC, 1, 0, NC, Ny, 0, 1, Ny, 0, 1.0, st);
		if (rc == BM_OK) rc = setup_line_fft(Ny, Tre, Aim, NC, NC, 1, 0, NC, Ny, 0, 1, Ny, 1, 1.0, st);
		if (rc == BM_OK) rc = setup_line_fft(Ny, Tim, Bre, NC, NC, 1, 0, NC, Ny, 0, 1, Ny, 0, 1.0, st);
		if (rc == BM_OK) rc = setup_line_fft(Ny, Tim, Bim, NC, NC, 1, 0, NC, Ny, 0, 1, Ny, 1, 1.0, st);
		if (rc) return rc;
		hipLaunchKernelGGL(k_cplx_combine, dim3((unsigned)std::min<size_t>((nk + 255) / 256, 4096)), dim3(256), 0, st,
			Are, Aim, Bre, Bim, Kp + (2 * q) * nk, Kp + (2 * q + 1) * nk, nk);
	}
	return BM_OK;
}

int compute_ml_kernels(bm_mlayer_s* m, hipStream_t st)
{
	const int L = m->L, Nx = m->Nx, Ny = m->Ny, NC = m->NC, NYH = m->NYH;
	const size_t kq = (size_t)NC * NYH;
	double hmax = 0;
	for (int l = 0; l < L; l++) hmax = std::max(hmax, std::max(m->hx, std::max(m->hy, m->hz[l])));
	BM_HIP(hipMemsetAsync(m->K, 0, (size_t)m->kcount * sizeof(float), st));
	// own-layer kernels: the single-layer 2D kernels (xx, yy, zz, xy), one per slot
	std::vector<char> done(m->nslot, 0);
	for (int l = 0; l < L; l++) {
		const int slot = m->hkidx[(size_t)l * L + l];
		if (done[slot]) continue;
		done[slot] = 1;
		bm_demag_t dl = nullptr;
		int rc = m->periodic() ? bm_demag_create_pbc(m->nx, m->ny, 1, m->hx, m->hy, m->hz[l], m->pbc[0], m->pbc[1], 0, 0, &dl)
		                       : bm_demag_create(m->nx, m->ny, 1, m->hx, m->hy, m->hz[l], 0, &dl);
		if (rc) return rc;
		float* dst = m->K + m->kbase(l, l) * 6 * kq;
		hipError_t e = hipMemcpyAsync(dst, dl->K, 4 * kq * sizeof(float), hipMemcpyDeviceToDevice, st);
		if (e == hipSuccess) e = hipStreamSynchronize(st);
		bm_demag_destroy(dl);
		if (e != hipSuccess) return hip_fail(e, "multilayer self kernels");
	}
	if (L == 1) return BM_OK;
	const int dx = m->pbc[0] ? kAsymDist : std::min(m->nx, kAsymDist), dy = m->pbc[1] ? kAsymDist : std::min(m->ny, kAsymDist);
	size_t gmax = (size_t)(std::max(dx, dy) + 2) * (std::max(dx, dy) + 2) * 3;
	// complex pairs: centred grids over at most 2 * kAsymDist + 5 offsets per axis, six NC x Ny temporaries
	if (m->Kc) gmax = std::max(gmax, (size_t)(2 * kAsymDist + 7) * (2 * kAsymDist + 7) * 3);
	const size_t tlen = (size_t)NC * Ny * (m->Kc ? 6 : 1);
	double* D = nullptr; double* T1 = nullptr; double* grids[12] = { nullptr };
	int rc = BM_OK;
	bool ok = hipMalloc(&D, (size_t)Nx * Ny * 6 * sizeof(double)) == hipSuccess && hipMalloc(&T1, tlen * sizeof(double)) == hipSuccess;
	for (int q = 0; q < 12 && ok; q++) ok = hipMalloc(&grids[q], gmax * sizeof(double)) == hipSuccess;
	if (!ok) rc = fail(BM_ERR_NOMEM, "multilayer kernel setup");
	for (int d = 0; d < L && rc == BM_OK; d++)
		for (int s = 0; s < L && rc == BM_OK; s++) {
			const int slot = m->hkidx[(size_t)d * L + s];
			if (d == s || done[slot]) continue;
			done[slot] = 1;
			const long long co = m->hkcoff.empty() ? -1 : m->hkcoff[(size_t)d * L + s];
			if (co >= 0) rc = shift_kernels(m, d, s, hmax, D, T1, grids, m->Kc + co, st);
			else rc = zshift_kernels(m, d, s, hmax, D, T1, grids, m->K + m->kbase(d, s) * 6 * kq, st);
		}
	// per-pair storage: pairs that share a slot take a copy of its first pair's kernel
	if (rc == BM_OK && m->per_pair()) {
		std::vector<int> first(m->nslot, -1);
		for (int pr = 0; pr < L * L; pr++) {
			const int slot = m->hkidx[pr];
			if (first[slot] < 0) { first[slot] = pr; continue; }
			const hipError_t e = hipMemcpyAsync(m->K + (size_t)pr * 6 * kq, m->K + (size_t)first[slot] * 6 * kq, 6 * kq * sizeof(float),
				hipMemcpyDeviceToDevice, st);
			if (e != hipSuccess) { rc = hip_fail(e, "multilayer kernel copies"); break; }
		}
	}
	if (rc == BM_OK) { hipError_t e = hipStreamSynchronize(st); if (e != hipSuccess) rc = hip_fail(e, "multilayer kernel setup"); }
	(void)hipStreamSynchronize(st);
	if (D) (void)hipFree(D);
	if (T1) (void)hipFree(T1);
	for (int q = 0; q < 12; q++) if (grids[q]) (void)hipFree(grids[q]);
	return rc;
}

} // namespace

using bm::fail;

// Handle on one grid of nx x ny cells with planned lengths (nx and 2 ny; along a periodic
// axis N = n: nx / 2 and ny).
static int ml_create_grid(int nlayers, int nx, int ny, double hx, double hy, const double* hz, const double* z0, bm_mlayer_t* out,
	int px = 0, int py = 0, const double* rx = nullptr, const double* ry = nullptr)
{
	bm_mlayer_s* m = new bm_mlayer_s();
	m->L = nlayers; m->nx = nx; m->ny = ny;
	m->pbc[0] = px; m->pbc[1] = py;
	m->Nx = px ? nx : 2 * nx; m->Ny = py ? ny : 2 * ny; m->NC = m->Nx / 2 + 1; m->NYH = m->Ny / 2 + 1; m->Sp = m->Ny;
	m->xs = m->Nx / 2;
	m->hx = hx; m->hy = hy; m->hz.assign(hz, hz + nlayers); m->z0.assign(z0, z0 + nlayers);
	m->ncu = bm::device_cu_count();
	m->ntl = bm::sx_tiles(ny);
	const size_t sbytes = (size_t)3 * nlayers * m->NC * m->Sp * sizeof(float2);
	const size_t sxbytes = (size_t)3 * nlayers * m->ntl * m->NC * bm::kLT * sizeof(float2);
	// layers offset in-plane against each other: those pairs take complex kernels
	long long ncx = 0;
	if (rx && ry) {
		m->rx.assign(rx, rx + nlayers); m->ry.assign(ry, ry + nlayers);
		m->hkcoff.assign((size_t)nlayers * nlayers, -1);
		for (int d = 0; d < nlayers; d++)
			for (int s = 0; s < nlayers; s++)
				if (d != s && (std::fabs(rx[d] - rx[s]) > 1e-12 * hx || std::fabs(ry[d] - ry[s]) > 1e-12 * hy))
					m->hkcoff[(size_t)d * nlayers + s] = (ncx++) * 12LL * m->NC * m->Ny;
		if (!ncx) { m->rx.clear(); m->ry.clear(); m->hkcoff.clear(); }
	}
	// kernel slots: a pair's kernel depends on the two thicknesses and the z shift only (the
	// lateral cells are common), so pairs that agree in them (to 1e-12 relative) share one;
	// a pair with a complex kernel has a slot of its own
	m->hkidx.assign((size_t)nlayers * nlayers, -1);
	struct Key { int self; double hd, hs, sh; };
	std::vector<Key> keys;
	auto same = [](double a, double b) { return std::fabs(a - b) <= 1e-12 * std::max(std::fabs(a), std::fabs(b)); };
	for (int d = 0; d < nlayers; d++)
		for (int s = 0; s < nlayers; s++) {
			const Key k{ d == s, hz[d], hz[s], d == s ? 0.0 : z0[d] - z0[s] };
			int slot = -1;
			const bool cplx = !m->hkcoff.empty() && m->hkcoff[(size_t)d * nlayers + s] >= 0;
			for (size_t q = 0; q < keys.size() && slot < 0 && !cplx; q++)
				if (keys[q].self == k.self && same(keys[q].hd, k.hd) && same(keys[q].hs, k.hs) && same(keys[q].sh, k.sh)) slot = (int)q;
			if (slot < 0) { slot = (int)keys.size(); keys.push_back(k); }
			m->hkidx[(size_t)d * nlayers + s] = slot;
		}
	m->nslot = (int)keys.size();
	m->kcount = (long long)(m->per_pair() ? nlayers * nlayers : m->nslot) * 6 * m->NC * m->NYH;
	if (hipMalloc(&m->Sx, sxbytes) != hipSuccess || hipMalloc(&m->S, sbytes) != hipSuccess
		|| ((nlayers > kRegLayers || ncx) && hipMalloc(&m->S2, sbytes) != hipSuccess)
		|| (ncx && hipMalloc(&m->Kc, (size_t)ncx * 12 * m->NC * m->Ny * sizeof(float)) != hipSuccess)
		|| (ncx && hipMalloc(&m->kcoff, m->hkcoff.size() * sizeof(long long)) != hipSuccess)
		|| hipMalloc(&m->K, (size_t)m->kcount * sizeof(float)) != hipSuccess
		|| (nlayers > kRegLayers && hipMalloc(&m->kidx, m->hkidx.size() * sizeof(int)) != hipSuccess)
		|| hipMalloc(&m->ebuf, (size_t)bm::x_plane_parts(m->xs, ny) * sizeof(double)) != hipSuccess) {
		bm_mlayer_destroy(m);
		return fail(BM_ERR_NOMEM, "bm_mlayer_create: device memory");
	}
	if ((m->kidx && hipMemcpy(m->kidx, m->hkidx.data(), m->hkidx.size() * sizeof(int), hipMemcpyHostToDevice) != hipSuccess)
		|| (m->kcoff && hipMemcpy(m->kcoff, m->hkcoff.data(), m->hkcoff.size() * sizeof(long long), hipMemcpyHostToDevice) != hipSuccess)) {
		bm_mlayer_destroy(m);
		return fail(BM_ERR_HIP, "bm_mlayer_create: pair table");
	}
	m->bytes = (long long)(sbytes * (m->S2 ? 2 : 1) + sxbytes) + m->kcount * 4 + ncx * 48LL * m->NC * m->Ny;
	int rc = compute_ml_kernels(m, nullptr);
	if (rc) { bm_mlayer_destroy(m); return rc; }
	*out = m;
	return BM_OK;
}

// Rectangle (m) for the mesh transfer tables.
struct R3 { double sx, sy, sz, ex, ey, ez; };

// The cells of a grid (origin o, cell h, nx x ny x 1 cells) that rectangle r covers, with the
// reference's reciprocal-distance weights d_max - |centre(r) - centre(cell)|, d_max =
// |size(r)/2 + h/2|, in its loop order (x, then y, then z), appended to out; returns their
// total.  Transfer::build_meshcells_weights / build_supermeshcells_weights
// (Boris/BorisLib/VEC_MeshTransfer.h:304-350, 396-440); the corner cell indices are its
// floor / ceil of a DBL3, which round with a 1e-4 cell tolerance (floor_epsilon,
// ceil_epsilon, Boris/BorisLib/Funcs_Math.h:161-173).
static double recip_weights(const R3& r, const double o[3], const double h[3], int nx, int ny, std::vector<std::pair<int, double>>& out)
{
	const double cx = (r.ex + r.sx) / 2 - o[0], cy = (r.ey + r.sy) / 2 - o[1], cz = (r.ez + r.sz) / 2 - o[2];
	const double eps = 1e-4;
	const int ill = (int)std::floor((r.sx - o[0]) / h[0] + eps), jll = (int)std::floor((r.sy - o[1]) / h[1] + eps), kll = (int)std::floor((r.sz - o[2]) / h[2] + eps);
	const int iur = (int)std::ceil((r.ex - o[0]) / h[0] - eps), jur = (int)std::ceil((r.ey - o[1]) / h[1] - eps), kur = (int)std::ceil((r.ez - o[2]) / h[2] - eps);
	const double mx = (r.ex - r.sx) / 2 + h[0] / 2, my = (r.ey - r.sy) / 2 + h[1] / 2, mz = (r.ez - r.sz) / 2 + h[2] / 2;
	const double dmax = std::sqrt(mx * mx + my * my + mz * mz);
	double tot = 0;
	for (int i = std::max(ill, 0); i < std::min(iur, nx); i++)
		for (int j = std::max(jll, 0); j < std::min(jur, ny); j++)
			for (int k = std::max(kll, 0); k < std::min(kur, 1); k++) {
				const double dx = cx - (i + 0.5) * h[0], dy = cy - (j + 0.5) * h[1], dz = cz - (k + 0.5) * h[2];
				const double d = dmax - std::sqrt(dx * dx + dy * dy + dz * dz);
				tot += d;
				out.push_back({ i + j * nx, d });
			}
	return tot;
}

// Covered volume ratio of cell c by rectangle m, 0 when they do not intersect in a volume (the
// reference's Rect::get_intersection / intersects / IsPlane / IsPoint, with its 1e-12
// absolute tolerance, Boris/BorisLib/Types_Rect.h:130-139, 347-361, 494-505).
static double covered_ratio(const R3& c, const R3& m)
{
	const double tol = 1e-12;
	const double sx = std::max(c.sx, m.sx), sy = std::max(c.sy, m.sy), sz = std::max(c.sz, m.sz);
	const double ex = std::min(c.ex, m.ex), ey = std::min(c.ey, m.ey), ez = std::min(c.ez, m.ez);
	auto zop = [&](double v) { return v > 0 || (v < tol && v > -tol); };
	if (!(zop(ex - sx) && zop(ey - sy) && zop(ez - sz))) return 0.0;
	auto isz = [&](double v) { return v < tol && v > -tol; };
	const double dx = ex - sx, dy = ey - sy, dz = ez - sz;
	const bool plane = (isz(dx) && !isz(dy) && !isz(dz)) || (isz(dy) && !isz(dx) && !isz(dz)) || (isz(dz) && !isz(dx) && !isz(dy));
	if (plane || (dx == 0 && dy == 0 && dz == 0)) return 0.0;
	return (dx / (c.ex - c.sx)) * (dy / (c.ey - c.sy)) * (dz / (c.ez - c.sz));
}

// Layer meshes of their own sizes, cells and positions (2D layers) on one common grid.  As
// the reference sets up the multilayered convolution (SDemag::set_default_n_common and
// set_Rect_collection, Boris/SDemag.cpp:173-300): n_common = the largest cell counts; every
// layer's rectangle is enlarged to the largest extents, starting as close to the smallest
// start as its end allows; its common cell is that rectangle over n_common.  The transfer of
// each layer to and from its rectangle's common cells is the reference's
// (SDemag_Demag::Initialize_Mesh_Transfer, Boris/SDemag_Demag.cpp:64-160): in, weighted by
// reciprocal distance and the covered volume ratio (MESHTRANSFERTYPE_WEIGHTED,
// Boris/BorisLib/VEC_MeshTransfer.h:899-945); out, by reciprocal distance (:1221-1260).  The
// reference couples rectangles at different places by shifted kernels; here every rectangle
// sits at its place in the rectangles' bounding box at the common cell (zero elsewhere), and
// the z-shifted kernels of that box give the same sums over the same cells — so the
// rectangles must lie on one grid of common cells (else BM_ERR_SIZE).  The box is padded with
// empty cells to planned FFT lengths (nx, 2 ny).
static int ml_create_meshes(int nlayers, const bm_layer_mesh* ms, bm_mlayer_t* out)
{
	double minx = ms[0].x0, miny = ms[0].y0, mxs = 0, mys = 0;
	int ncx = 1, ncy = 1;
	for (int l = 0; l < nlayers; l++) {
		const bm_layer_mesh& q = ms[l];
		if (q.nx < 1 || q.ny < 1 || !(q.hx > 0) || !(q.hy > 0) || !(q.hz > 0)) return fail(BM_ERR_ARG, "bm_mlayer_create_meshes: bad layer mesh");
		ncx = std::max(ncx, q.nx); ncy = std::max(ncy, q.ny);
		mxs = std::max(mxs, q.nx * q.hx); mys = std::max(mys, q.ny * q.hy);
		minx = std::min(minx, q.x0); miny = std::min(miny, q.y0);
	}
	// enlarged rectangles
	std::vector<R3> Rm((size_t)nlayers), Rc((size_t)nlayers);
	for (int l = 0; l < nlayers; l++) {
		const bm_layer_mesh& q = ms[l];
		R3& m = Rm[(size_t)l];
		m = R3{ q.x0, q.y0, q.z0, q.x0 + q.nx * q.hx, q.y0 + q.ny * q.hy, q.z0 + q.hz };
		R3 c = m;
		if (c.ex - mxs < minx) { c.sx = minx; c.ex = c.sx + mxs; } else c.sx = c.ex - mxs;
		if (c.ey - mys < miny) { c.sy = miny; c.ey = c.sy + mys; } else c.sy = c.ey - mys;
		Rc[(size_t)l] = c;
	}
	// the bounding box of the rectangles at the common cell
	const double hcx = mxs / ncx, hcy = mys / ncy;
	double bx = Rc[0].sx, by = Rc[0].sy, bex = Rc[0].ex, bey = Rc[0].ey;
	for (const R3& c : Rc) { bx = std::min(bx, c.sx); by = std::min(by, c.sy); bex = std::max(bex, c.ex); bey = std::max(bey, c.ey); }
	(void)bex; (void)bey;
	// each rectangle starts a whole number of common cells into the box plus a fraction of one:
	// the fractions couple the layers through kernels shifted in-plane (complex kernels)
	auto split = [](double v, int& n, double& r) { const double f = std::floor(v + 1e-6); n = (int)f; r = std::fabs(v - f) <= 1e-6 ? 0.0 : v - f; };
	int nbx = 0, nby = 0;
	std::vector<int> ox((size_t)nlayers), oy((size_t)nlayers);
	std::vector<double> rxs((size_t)nlayers), rys((size_t)nlayers);
	bool frac = false;
	for (int l = 0; l < nlayers; l++) {
		double fx, fy;
		split((Rc[(size_t)l].sx - bx) / hcx, ox[(size_t)l], fx);
		split((Rc[(size_t)l].sy - by) / hcy, oy[(size_t)l], fy);
		rxs[(size_t)l] = fx * hcx; rys[(size_t)l] = fy * hcy;
		frac = frac || fx != 0.0 || fy != 0.0;
		nbx = std::max(nbx, ox[(size_t)l] + ncx); nby = std::max(nby, oy[(size_t)l] + ncy);
	}
	const int nxp = plan_length_at_least(std::max(2, nbx)), nyp = plan_length_at_least(2 * nby) / 2;
	if (!nxp || nyp < 1 || !bm::plan_2d_ok(nxp, 2 * nyp)) return fail(BM_ERR_SIZE, "bm_mlayer_create_meshes: no FFT plan for the common grid");
	// the layers' own grids are the common one: nothing to transfer
	bool plain = nxp == nbx && nyp == nby;
	for (int l = 0; l < nlayers; l++) {
		const bm_layer_mesh& q = ms[l];
		plain = plain && q.nx == nbx && q.ny == nby && ox[(size_t)l] == 0 && oy[(size_t)l] == 0
			&& std::fabs(q.hx - hcx) <= 1e-12 * hcx && std::fabs(q.hy - hcy) <= 1e-12 * hcy;
	}
	std::vector<double> hz((size_t)nlayers), z0((size_t)nlayers);
	for (int l = 0; l < nlayers; l++) { hz[(size_t)l] = ms[l].hz; z0[(size_t)l] = ms[l].z0; }
	bm_mlayer_t m = nullptr;
	int rc = ml_create_grid(nlayers, nxp, nyp, hcx, hcy, hz.data(), z0.data(), &m, 0, 0, frac ? rxs.data() : nullptr, frac ? rys.data() : nullptr);
	if (rc || plain) { if (!rc) *out = m; return rc; }
	// transfer tables
	const size_t plane = (size_t)nxp * nyp;
	std::vector<long long> moff((size_t)nlayers + 1, 0);
	for (int l = 0; l < nlayers; l++) moff[(size_t)l + 1] = moff[(size_t)l] + (long long)ms[l].nx * ms[l].ny;
	std::vector<int> iptr(1, 0), icol, optr(1, 0), ocol;
	std::vector<double> ival, oval;
	std::vector<std::pair<int, double>> w;
	for (int l = 0; l < nlayers; l++) {
		const bm_layer_mesh& q = ms[l];
		const R3& c = Rc[(size_t)l];
		const R3& mr = Rm[(size_t)l];
		const double hc[3] = { (c.ex - c.sx) / ncx, (c.ey - c.sy) / ncy, c.ez - c.sz };
		const double hm[3] = { q.hx, q.hy, q.hz };
		const double om[3] = { mr.sx, mr.sy, mr.sz }, oc[3] = { c.sx, c.sy, c.sz };
		// in: every cell of the layer's plane of the box; cells of its rectangle take the layer cells
		for (size_t cell = 0; cell < plane; cell++) {
			const int bi = (int)(cell % nxp), bj = (int)(cell / nxp), i = bi - ox[(size_t)l], j = bj - oy[(size_t)l];
			if (i >= 0 && j >= 0 && i < ncx && j < ncy) {
				const R3 rc3{ c.sx + i * hc[0], c.sy + j * hc[1], c.sz, c.sx + i * hc[0] + hc[0], c.sy + j * hc[1] + hc[1], c.sz + hc[2] };
				const double cov = covered_ratio(rc3, mr);
				if (cov > 0) {
					w.clear();
					const double tot = recip_weights(rc3, om, hm, q.nx, q.ny, w);
					for (auto& e : w) { icol.push_back((int)(moff[(size_t)l] + e.first)); ival.push_back(tot > 0 ? e.second * (cov / tot) : e.second); }
				}
			}
			iptr.push_back((int)icol.size());
		}
		// out: every layer cell from the common cells of its rectangle (box plane indices)
		for (int J = 0; J < q.ny; J++)
			for (int I = 0; I < q.nx; I++) {
				const R3 r3{ mr.sx + I * hm[0], mr.sy + J * hm[1], mr.sz, mr.sx + I * hm[0] + hm[0], mr.sy + J * hm[1] + hm[1], mr.sz + hm[2] };
				w.clear();
				const double tot = recip_weights(r3, oc, hc, ncx, ncy, w);
				for (auto& e : w) {
					const int i = e.first % ncx, j = e.first / ncx;
					ocol.push_back((oy[(size_t)l] + j) * nxp + ox[(size_t)l] + i);
					oval.push_back(tot > 0 ? e.second * (1.0 / tot) : e.second);
				}
				optr.push_back((int)ocol.size());
			}
	}
	const size_t nb = (iptr.size() + optr.size() + icol.size() + ocol.size()) * sizeof(int) + (ival.size() + oval.size()) * sizeof(double) + 64;
	char* t = nullptr;
	if (hipMalloc(&t, nb) != hipSuccess || hipMalloc(&m->stage, 3 * plane * (size_t)(nlayers + 1) * sizeof(float)) != hipSuccess
		|| hipMalloc(&m->epart, 1024 * sizeof(double)) != hipSuccess) {
		if (t) (void)hipFree(t);
		bm_mlayer_destroy(m);
		return fail(BM_ERR_NOMEM, "bm_mlayer_create_meshes: device memory");
	}
	m->tabs = t;
	size_t at = 0;
	auto put = [&](const void* src, size_t bytes) { void* d = t + at; (void)hipMemcpy(d, src, bytes, hipMemcpyHostToDevice); at += (bytes + 7) / 8 * 8; return d; };
	m->tin.val = (const double*)put(ival.data(), ival.size() * sizeof(double));
	m->tout.val = (const double*)put(oval.data(), oval.size() * sizeof(double));
	m->tin.ptr = (const int*)put(iptr.data(), iptr.size() * sizeof(int));
	m->tin.col = (const int*)put(icol.data(), icol.size() * sizeof(int));
	m->tout.ptr = (const int*)put(optr.data(), optr.size() * sizeof(int));
	m->tout.col = (const int*)put(ocol.data(), ocol.size() * sizeof(int));
	m->moff = moff;
	m->bytes += (long long)(nb + 3 * plane * (nlayers + 1) * sizeof(float));
	*out = m;
	return BM_OK;
}

static int ml_check_layers(int nlayers, const double* hz)
{
	if (nlayers > kMaxLayers) return fail(BM_ERR_SIZE, "bm_mlayer_create: at most 64 layers");
	for (int l = 0; l < nlayers; l++) if (!(hz[l] > 0)) return fail(BM_ERR_ARG, "bm_mlayer_create: layer thickness must be positive");
	return BM_OK;
}

extern "C" int bm_mlayer_create(int nlayers, int nx, int ny, double hx, double hy, const double* hz, const double* z0, int flags, bm_mlayer_t* out)
{
	(void)flags;
	if (!out || !hz || !z0 || nlayers < 1 || nx < 1 || ny < 1 || !(hx > 0) || !(hy > 0)) return fail(BM_ERR_ARG, "bm_mlayer_create: bad arguments");
	*out = nullptr;
	int rc = ml_check_layers(nlayers, hz);
	if (rc) return rc;
	if (nx % 2 == 0 && bm::plan_2d_ok(nx, 2 * ny)) return ml_create_grid(nlayers, nx, ny, hx, hy, hz, z0, out);
	// sizes without a plan: the layers on a padded common grid
	std::vector<bm_layer_mesh> ms((size_t)nlayers);
	for (int l = 0; l < nlayers; l++) ms[(size_t)l] = bm_layer_mesh{ nx, ny, hx, hy, hz[l], 0.0, 0.0, z0[l] };
	return ml_create_meshes(nlayers, ms.data(), out);
}

// Periodic images along x and y (SDemag::Set_PBC, Boris/SDemag.cpp:403-408): N = n along a
// periodic axis (nx even with plans for nx / 2 and nx, ny with a plan), the self kernels are
// the periodic 2D kernels and the pair kernels the periodic z-shifted ones; layers of one grid
// only (no images along z).
extern "C" int bm_mlayer_create_pbc(int nlayers, int nx, int ny, double hx, double hy, const double* hz, const double* z0,
	int pbc_x, int pbc_y, int flags, bm_mlayer_t* out)
{
	if (pbc_x < 0 || pbc_y < 0) return fail(BM_ERR_ARG, "bm_mlayer_create_pbc: negative image count");
	if (!pbc_x && !pbc_y) return bm_mlayer_create(nlayers, nx, ny, hx, hy, hz, z0, flags, out);
	if (!out || !hz || !z0 || nlayers < 1 || nx < 1 || ny < 1 || !(hx > 0) || !(hy > 0)) return fail(BM_ERR_ARG, "bm_mlayer_create_pbc: bad arguments");
	*out = nullptr;
	int rc = ml_check_layers(nlayers, hz);
	if (rc) return rc;
	if (!bm::plan_2d_pbc_ok(nx, ny, pbc_x > 0, pbc_y > 0))
		return fail(BM_ERR_SIZE, "bm_mlayer_create_pbc: no FFT plan (a periodic nx: even, with plans for nx / 2 and nx; a periodic ny: a plan "
			"for ny; open axes: plans for nx and 2 ny)");
	return ml_create_grid(nlayers, nx, ny, hx, hy, hz, z0, out, pbc_x, pbc_y);
}

void bm::mlayer_pbc(const bm_mlayer_s* m, int out[2]) { out[0] = m->pbc[0]; out[1] = m->pbc[1]; }

extern "C" int bm_mlayer_create_meshes(int nlayers, const bm_layer_mesh* meshes, int flags, bm_mlayer_t* out)
{
	(void)flags;
	if (!out || !meshes || nlayers < 1) return fail(BM_ERR_ARG, "bm_mlayer_create_meshes: bad arguments");
	*out = nullptr;
	std::vector<double> hz((size_t)nlayers);
	for (int l = 0; l < nlayers; l++) hz[(size_t)l] = meshes[l].hz;
	int rc = ml_check_layers(nlayers, hz.data());
	return rc ? rc : ml_create_meshes(nlayers, meshes, out);
}

extern "C" int bm_mlayer_destroy(bm_mlayer_t m)
{
	if (!m) return BM_OK;
	if (m->Sx) (void)hipFree(m->Sx);
	if (m->S) (void)hipFree(m->S);
	if (m->S2) (void)hipFree(m->S2);
	if (m->K) (void)hipFree(m->K);
	if (m->Kc) (void)hipFree(m->Kc);
	if (m->kcoff) (void)hipFree(m->kcoff);
	if (m->kidx) (void)hipFree(m->kidx);
	if (m->ebuf) (void)hipFree(m->ebuf);
	if (m->tabs) (void)hipFree(m->tabs);
	if (m->stage) (void)hipFree(m->stage);
	if (m->epart) (void)hipFree(m->epart);
	delete m;
	return BM_OK;
}

// The convolution (tr: its transpose) into out with the final x C2R in `mode` (0 planar
// set, 1 AoS set, 2 AoS add).
static int ml_convolute(bm_mlayer_s* m, const float* M, float* out, int mode, hipStream_t st, int tr = 0)
{
	const int L = m->L, nlines = 3 * L * m->NC;
	const bool fx = m->pbc[0] > 0, fy = m->pbc[1] > 0;
	int rc = bm::pass_fwd_x(m->xs, M, m->Sx, m->ny, L, m->ntl, m->ncu, st, fx);
	if (rc == BM_OK) rc = bm::pass_y(m->Ny, -1, m->S, m->Sx, nlines, m->Sp, m->NC, m->ntl, m->ncu, st, fy);
	if (rc == BM_OK) {
		bm::prof_mark_begin(bm::PK_ML_MULT, st);
		rc = launch_ml_mult(L, m->S, m->S2, m->K, m->kidx, m->NC, m->Ny, m->Sp, m->ncu, st, tr, m->Kc, m->kcoff);
		bm::prof_mark_end(bm::PK_ML_MULT, st);
	}
	if (rc == BM_OK) rc = bm::pass_y(m->Ny, 1, m->Sout(), m->Sx, nlines, m->Sp, m->NC, m->ntl, m->ncu, st, fy);
	const float scale = (float)(1.0 / ((double)m->Nx * m->Ny));
	if (rc == BM_OK) rc = bm::pass_inv_x(m->xs, mode, m->Sx, out, m->ny, L, m->ntl, scale, m->ncu, st, fx);
	return rc;
}

int bm::mlayer_planar(bm_mlayer_s* m, const float* M, float* H3n, hipStream_t st) { return ml_convolute(m, M, H3n, 0, st); }
int bm::mlayer_planar_transposed(bm_mlayer_s* m, const float* M, float* H3n, hipStream_t st) { return ml_convolute(m, M, H3n, 0, st, 1); }

void bm::mlayer_cell(const bm_mlayer_s* m, int l, double h[3])
{
	h[0] = m->hx; h[1] = m->hy; h[2] = m->hz[(size_t)l];
}

double bm::mlayer_z0(const bm_mlayer_s* m, int l) { return m->z0[(size_t)l]; }

bm::MlView bm::mlayer_view(bm_mlayer_s* m)
{
	return MlView{ m->Sx, m->S, m->K, m->L, m->nx, m->ny, m->Nx, m->Ny, m->NC, m->NYH, m->Sp, m->ntl };
}

int bm::batch_ml_mult(int L, LaunchShape sh, const MlBatch* h_tab, const MlBatch* d_tab, int nb, int ncu, hipStream_t st)
{
	int mx = 1;
	for (int b = 0; b < nb; b++) mx = std::max(mx, h_tab[b].NC * h_tab[b].Ny);
	dim3 g((unsigned)std::max(1, std::min((mx + 255) / 256, ncu * (sh.per_cu ? sh.per_cu : 8))), nb), bl(256);
	switch (L) {
#define C_(V_) case V_: hipLaunchKernelGGL((k_ml_mult<V_>), g, bl, 0, st, nullptr, nullptr, 0, 0, 0, 0, d_tab); return BM_OK;
	C_(1) C_(2) C_(3) C_(4) C_(5) C_(6) C_(7) C_(8) C_(9) C_(10) C_(11) C_(12) C_(13) C_(14) C_(15) C_(16)
#undef C_
	}
	return fail(BM_ERR_SIZE, "batched multilayer demag: at most 16 layers per stack");
}

bool bm::mlayer_mapped(const bm_mlayer_s* m) { return m->mapped(); }

void bm::mlayer_dims(const bm_mlayer_s* m, int* L, int* nx, int* ny, double* hz)
{
	*L = m->L; *nx = m->nx; *ny = m->ny;
	if (hz) for (int l = 0; l < m->L; l++) hz[l] = m->hz[l];
}

static int ml_grid(long long n) { return (int)std::max<long long>(1, std::min<long long>((n + 255) / 256, 8192)); }

extern "C" int bm_mlayer_convolute(bm_mlayer_t m, const float* M, float* Heff, int clear_out, void* stream)
{
	if (!m || !M || !Heff) return fail(BM_ERR_ARG, "bm_mlayer_convolute: null argument");
	hipStream_t st = (hipStream_t)stream;
	int rc;
	if (m->mapped()) {
		// layer meshes -> common planes -> convolution (fields over the same planes) -> meshes
		const size_t plane = (size_t)m->nx * m->ny;
		const long long rows = (long long)plane * m->L;
		hipLaunchKernelGGL(k_ml_transfer_in, dim3(ml_grid(rows)), dim3(256), 0, st, m->tin, M, m->stage, 0LL, rows, 0LL);
		rc = ml_convolute(m, m->stage, m->stage, 1, st);
		// each layer's cells from its own plane
		for (int l = 0; l < m->L && rc == BM_OK; l++) {
			const long long r0 = m->moff[(size_t)l], nr = m->moff[(size_t)l + 1] - r0;
			hipLaunchKernelGGL(k_ml_transfer_out, dim3(ml_grid(nr)), dim3(256), 0, st, m->tout, (const float*)(m->stage + 3 * plane * (size_t)l),
				Heff + 3 * r0, r0, nr, clear_out ? 1 : 2, (const float*)nullptr, (double*)nullptr);
		}
	}
	else rc = ml_convolute(m, M, Heff, clear_out ? 1 : 2, st);
	if (rc) return rc;
	BM_HIP(hipGetLastError());
	return BM_OK;
}

extern "C" int bm_mlayer_copy_kernels(bm_mlayer_t m, float* host_out, long long count)
{
	const size_t per = (size_t)6 * m->NC * m->NYH;
	if (!m || !host_out || count < (long long)((size_t)m->L * m->L * per)) return fail(BM_ERR_ARG, "bm_mlayer_copy_kernels: bad arguments");
	// pair by pair (d * L + s), each from its slot
	for (int d = 0; d < m->L; d++)
		for (int s = 0; s < m->L; s++)
			BM_HIP(hipMemcpy(host_out + ((size_t)d * m->L + s) * per, m->K + m->kbase(d, s) * per, per * sizeof(float), hipMemcpyDeviceToHost));
	return BM_OK;
}

// Split form, one call per layer mesh: the reference's ConvolutionCUDA::ForwardFFT,
// KernelMultiplication_MultipleInputs and InverseFFT (Boris/ConvolutionCUDA.h:254-351) as
// SDemagCUDA::UpdateField calls them for each layer's own mesh (Boris/SDemagCUDA.cpp:270+).
// A layer's forward transform fills its planes of the shared spectra (x pass into its Sx
// plane, then the y pass of its three component planes); the product runs once over all
// layers; the inverse of a layer reads back its planes.
extern "C" int bm_mlayer_forward(bm_mlayer_t m, int layer, const float* M_layer, void* stream)
{
	if (!m || !M_layer || layer < 0 || layer >= m->L) return fail(BM_ERR_ARG, "bm_mlayer_forward: bad arguments");
	hipStream_t st = (hipStream_t)stream;
	const float* src = M_layer;
	if (m->mapped()) {
		// the layer's mesh onto its common plane
		const size_t plane = (size_t)m->nx * m->ny;
		src = m->stage + 3 * plane * (size_t)layer;
		hipLaunchKernelGGL(k_ml_transfer_in, dim3(ml_grid((long long)plane)), dim3(256), 0, st, m->tin, M_layer, m->stage + 3 * plane * (size_t)layer,
			(long long)plane * layer, (long long)plane, m->moff[(size_t)layer]);
	}
	int rc = bm::pass_fwd_x_plane(m->xs, src, m->Sx, m->ny, layer, m->L, m->ntl, m->ncu, st, m->pbc[0] > 0);
	for (int c = 0; c < 3 && rc == BM_OK; c++) {
		const size_t cz = (size_t)c * m->L + layer;
		rc = bm::pass_y(m->Ny, -1, m->S + cz * m->NC * m->Sp, m->Sx + cz * m->ntl * m->NC * bm::kLT, m->NC, m->Sp, m->NC, m->ntl, m->ncu, st,
			m->pbc[1] > 0);
	}
	if (rc) return rc;
	BM_HIP(hipGetLastError());
	return BM_OK;
}

extern "C" int bm_mlayer_multiply(bm_mlayer_t m, void* stream)
{
	if (!m) return fail(BM_ERR_ARG, "bm_mlayer_multiply: null handle");
	hipStream_t st = (hipStream_t)stream;
	bm::prof_mark_begin(bm::PK_ML_MULT, st);
	int rc = launch_ml_mult(m->L, m->S, m->S2, m->K, m->kidx, m->NC, m->Ny, m->Sp, m->ncu, st, 0, m->Kc, m->kcoff);
	bm::prof_mark_end(bm::PK_ML_MULT, st);
	if (rc) return rc;
	BM_HIP(hipGetLastError());
	return BM_OK;
}

extern "C" int bm_mlayer_inverse(bm_mlayer_t m, int layer, const float* M_layer, float* H_layer, int clear_out, float* energy, void* stream)
{
	if (!m || !H_layer || layer < 0 || layer >= m->L || (energy && !M_layer)) return fail(BM_ERR_ARG, "bm_mlayer_inverse: bad arguments");
	hipStream_t st = (hipStream_t)stream;
	int rc = BM_OK;
	for (int c = 0; c < 3 && rc == BM_OK; c++) {
		const size_t cz = (size_t)c * m->L + layer;
		rc = bm::pass_y(m->Ny, 1, m->Sout() + cz * m->NC * m->Sp, m->Sx + cz * m->ntl * m->NC * bm::kLT, m->NC, m->Sp, m->NC, m->ntl, m->ncu, st,
			m->pbc[1] > 0);
	}
	const float scale = (float)(1.0 / ((double)m->Nx * m->Ny));
	if (rc == BM_OK && m->mapped()) {
		// field on the common plane, then onto the layer's cells (and their energy density
		// from the layer's own M and the transferred field)
		const size_t plane = (size_t)m->nx * m->ny;
		float* Hp = m->stage + 3 * plane * (size_t)m->L;
		rc = bm::pass_inv_x_plane(m->xs, 1, m->Sx, Hp, m->ny, layer, m->L, m->ntl, scale, m->ncu, st, nullptr, nullptr, nullptr, m->pbc[0] > 0);
		const long long r0 = m->moff[(size_t)layer], nr = m->moff[(size_t)layer + 1] - r0;
		const int g = std::min(ml_grid(nr), 1024);
		if (rc == BM_OK) {
			hipLaunchKernelGGL(k_ml_transfer_out, dim3(g), dim3(256), 0, st, m->tout, (const float*)Hp, H_layer, r0, nr, clear_out ? 1 : 2,
				energy ? M_layer : (const float*)nullptr, m->epart);
			if (energy) hipLaunchKernelGGL(k_ml_energy, dim3(1), dim3(64), 0, st, (const double*)m->epart, g, nr, energy);
		}
	}
	else if (rc == BM_OK)
		rc = bm::pass_inv_x_plane(m->xs, clear_out ? 1 : 2, m->Sx, H_layer, m->ny, layer, m->L, m->ntl, scale, m->ncu, st, M_layer, m->ebuf, energy,
			m->pbc[0] > 0);
	if (rc) return rc;
	BM_HIP(hipGetLastError());
	return BM_OK;
}
