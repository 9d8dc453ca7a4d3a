// ORACLE — TEST INFRASTRUCTURE ONLY.
// CPU FP64 restatement of the reference's multilayered (stacked 2D layers) demag: the
// z-shifted source→destination tensors, regular and irregular (source and destination
// cells differing in thickness), their real kernels, and the multi-input kernel product.
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load or call
// anything under oracle/; the product (boris2_amd/) never does.
//
// Parity status (pins in tests/test_oracle_mlayer.py):
//   tensor2d_zshift (regular and irregular, near and far field): bit-for-bit against
//       oracle/_ref (the reference's Boris/DemagTFunc_Shifted.cpp + DemagTFunc_Irregular.cpp
//       compiled here) through committed golden vectors; where no far field is involved,
//       also against geometry: a source layer of thickness 2h equals two stacked layers
//       of thickness h, a destination of thickness 2h averages its two halves, and the
//       regular z-shifted tensor of touching layers equals the 3D tensor's kz = 1 plane.
//   kernels2d_zshift, MLayer::convolute: against a direct real-space sum with the pinned
//       tensors, and an L-layer stack of equal touching layers against the 3D oracle
//       convolution of an nz = L mesh.
#pragma once

#include "oracle_tensor.hpp"
#include "oracle_core.hpp"

#include <cmath>
#include <complex>
#include <functional>
#include <vector>

namespace oracle {

// floor(x + 1e-4) (BorisLib/Funcs_Math.h:161, EPSILON_ROUNDING BorisLib/Funcs_Aux_base.h:13)
inline int floor_eps(double x) { return (int)std::floor(x + 1e-4); }

// Grid of a basis function at arbitrary per-axis positions: v(i, j, k) = fn(X(i), Y(j), Z(k))
// for i in [-1, dx], j in [-1, dy], k in [-1, dz], stored with offset +1 (x fastest).
inline Grid3 fill_grid_at(double (*fn)(double, double, double), int dx, int dy, int dz,
	const std::function<double(int)>& X, const std::function<double(int)>& Y, const std::function<double(int)>& Z)
{
	Grid3 g; g.nx = dx + 2; g.ny = dy + 2; g.nz_ = dz + 2;
	g.v.assign((size_t)g.nx * g.ny * g.nz_, 0.0);
	for (int k = -1; k <= dz; k++)
		for (int j = -1; j <= dy; j++)
			for (int i = -1; i <= dx; i++)
				g.v[(size_t)(i + 1) + (size_t)(j + 1) * g.nx + (size_t)(k + 1) * g.nx * g.ny] = fn(X(i), Y(j), Z(k));
	return g;
}

// ---------------------------------------------------------------------------------------
// 36-term second differences for irregular source/destination cells (thicknesses differ),
// Boris/DemagTFunc_L.cpp:596-830.  g holds the three z (or first-argument) planes
// -1, 0, +1 at the irregular positions, del the single extra plane; term order as the
// reference's (the Kahan-Neumaier sum is order-sensitive).  tau = destination volume.

// xx and yy (Ldia_zshifted_irregular_xx_yy) and xy (Lodia_zshifted_irregular_xy): the
// plane axis is the third one.
inline double L36_plane3(const Grid3& g, const Grid3& del, int i, int j, double tau)
{
	const int k = 1; i++; j++;
	auto G = [&](int a, int b, int c) { return g.at(a, b, c); };
	auto D = [&](int a, int b) { return del.at(a, b, 0); };
	double t[36];
	t[0] = +4 * G(i, j, k);          t[1] = +4 * D(i, j);
	t[2] = -2 * G(i + 1, j, k);      t[3] = -2 * D(i + 1, j);
	t[4] = -2 * G(i - 1, j, k);      t[5] = -2 * D(i - 1, j);
	t[6] = -2 * G(i, j + 1, k);      t[7] = -2 * D(i, j + 1);
	t[8] = -2 * G(i, j - 1, k);      t[9] = -2 * D(i, j - 1);
	t[10] = -4 * G(i, j, k + 1);     t[11] = -4 * G(i, j, k - 1);
	t[12] = +G(i + 1, j + 1, k);     t[13] = +D(i + 1, j + 1);
	t[14] = +G(i + 1, j - 1, k);     t[15] = +D(i + 1, j - 1);
	t[16] = +G(i - 1, j + 1, k);     t[17] = +D(i - 1, j + 1);
	t[18] = +G(i - 1, j - 1, k);     t[19] = +D(i - 1, j - 1);
	t[20] = +2 * G(i + 1, j, k + 1); t[21] = +2 * G(i + 1, j, k - 1);
	t[22] = +2 * G(i - 1, j, k + 1); t[23] = +2 * G(i - 1, j, k - 1);
	t[24] = +2 * G(i, j + 1, k + 1); t[25] = +2 * G(i, j + 1, k - 1);
	t[26] = +2 * G(i, j - 1, k + 1); t[27] = +2 * G(i, j - 1, k - 1);
	t[28] = -G(i - 1, j - 1, k - 1); t[29] = -G(i - 1, j - 1, k + 1);
	t[30] = -G(i - 1, j + 1, k - 1); t[31] = -G(i + 1, j - 1, k - 1);
	t[32] = -G(i - 1, j + 1, k + 1); t[33] = -G(i + 1, j - 1, k + 1);
	t[34] = -G(i + 1, j + 1, k - 1); t[35] = -G(i + 1, j + 1, k + 1);
	return kn_sum(t, 36) / (4 * kPI * tau);
}

// zz (Ldia_zshifted_irregular_zz): the plane axis is the first one; called at (0, j, i).
inline double L36_zz(const Grid3& g, const Grid3& del, int j, int k, double tau)
{
	const int i = 1; j++; k++;
	auto G = [&](int a, int b, int c) { return g.at(a, b, c); };
	auto D = [&](int b, int c) { return del.at(0, b, c); };
	double t[36];
	t[0] = +4 * G(i, j, k);          t[1] = +4 * D(j, k);
	t[2] = -4 * G(i + 1, j, k);      t[3] = -4 * G(i - 1, j, k);
	t[4] = -2 * G(i, j, k + 1);      t[5] = -2 * D(j, k + 1);
	t[6] = -2 * G(i, j, k - 1);      t[7] = -2 * D(j, k - 1);
	t[8] = -2 * G(i, j + 1, k);      t[9] = -2 * D(j + 1, k);
	t[10] = -2 * G(i, j - 1, k);     t[11] = -2 * D(j - 1, k);
	t[12] = +2 * G(i + 1, j + 1, k); t[13] = +2 * G(i + 1, j - 1, k);
	t[14] = +2 * G(i - 1, j + 1, k); t[15] = +2 * G(i - 1, j - 1, k);
	t[16] = +2 * G(i + 1, j, k + 1); t[17] = +2 * G(i + 1, j, k - 1);
	t[18] = +2 * G(i - 1, j, k + 1); t[19] = +2 * G(i - 1, j, k - 1);
	t[20] = +G(i, j + 1, k + 1);     t[21] = +D(j + 1, k + 1);
	t[22] = +G(i, j + 1, k - 1);     t[23] = +D(j + 1, k - 1);
	t[24] = +G(i, j - 1, k + 1);     t[25] = +D(j - 1, k + 1);
	t[26] = +G(i, j - 1, k - 1);     t[27] = +D(j - 1, k - 1);
	t[28] = -G(i - 1, j - 1, k - 1); t[29] = -G(i - 1, j - 1, k + 1);
	t[30] = -G(i - 1, j + 1, k - 1); t[31] = -G(i + 1, j - 1, k - 1);
	t[32] = -G(i - 1, j + 1, k + 1); t[33] = -G(i + 1, j - 1, k + 1);
	t[34] = -G(i + 1, j + 1, k - 1); t[35] = -G(i + 1, j + 1, k + 1);
	return kn_sum(t, 36) / (4 * kPI * tau);
}

// xz and yz (Lodia_zshifted_irregular_xz_yz): the plane axis is the second one; called at
// (i, 0, k).
inline double L36_plane2(const Grid3& g, const Grid3& del, int i, int k, double tau)
{
	const int j = 1; i++; k++;
	auto G = [&](int a, int b, int c) { return g.at(a, b, c); };
	auto D = [&](int a, int c) { return del.at(a, 0, c); };
	double t[36];
	t[0] = +4 * G(i, j, k);          t[1] = +4 * D(i, k);
	t[2] = -2 * G(i + 1, j, k);      t[3] = -2 * D(i + 1, k);
	t[4] = -2 * G(i - 1, j, k);      t[5] = -2 * D(i - 1, k);
	t[6] = -4 * G(i, j + 1, k);      t[7] = -4 * G(i, j - 1, k);
	t[8] = -2 * G(i, j, k + 1);      t[9] = -2 * D(i, k + 1);
	t[10] = -2 * G(i, j, k - 1);     t[11] = -2 * D(i, k - 1);
	t[12] = +2 * G(i + 1, j + 1, k); t[13] = +2 * G(i + 1, j - 1, k);
	t[14] = +2 * G(i - 1, j + 1, k); t[15] = +2 * G(i - 1, j - 1, k);
	t[16] = +G(i + 1, j, k + 1);     t[17] = +D(i + 1, k + 1);
	t[18] = +G(i + 1, j, k - 1);     t[19] = +D(i + 1, k - 1);
	t[20] = +G(i - 1, j, k + 1);     t[21] = +D(i - 1, k + 1);
	t[22] = +G(i - 1, j, k - 1);     t[23] = +D(i - 1, k - 1);
	t[24] = +2 * G(i, j + 1, k + 1); t[25] = +2 * G(i, j + 1, k - 1);
	t[26] = +2 * G(i, j - 1, k + 1); t[27] = +2 * G(i, j - 1, k - 1);
	t[28] = -G(i - 1, j - 1, k - 1); t[29] = -G(i - 1, j - 1, k + 1);
	t[30] = -G(i - 1, j + 1, k - 1); t[31] = -G(i + 1, j - 1, k - 1);
	t[32] = -G(i - 1, j + 1, k + 1); t[33] = -G(i + 1, j - 1, k + 1);
	t[34] = -G(i + 1, j + 1, k - 1); t[35] = -G(i + 1, j + 1, k + 1);
	return kn_sum(t, 36) / (4 * kPI * tau);
}

// Does DBL3 equality (component differences below 1e-12F, BorisLib/Types_VAL.h:500) hold
// for s*1e-9 vs d*1e-9, the test Boris/DemagTFunc_Irregular.cpp:11 uses to pick the
// regular path?
inline bool same_cell(const double s[3], const double d[3])
{
	for (int c = 0; c < 3; c++) if (nz(s[c] * 1e-9 - d[c] * 1e-9)) return false;
	return true;
}

// 2D z-shifted tensor between source cells s and destination cells d (normalised, equal
// in x and y) for a destination layer shifted by shz above the source layer, stored as
// -N on the full Nx x Ny grid (x fastest): Dd = (xx, yy, zz), Do = (xy, xz, yz).
// Regular (s == d): Boris/DemagTFunc_Shifted.cpp:235-449 with the z-shift grids of
// Boris/DemagTFunc_fg.cpp:350-410.  Irregular: Boris/DemagTFunc_Irregular.cpp:5-235 with
// Boris/DemagTFunc_fg.cpp:507-620; in the far field the equal-cell series of the
// destination cell is scaled by the source thickness s.z, as the reference does.
inline void tensor2d_zshift(int nx, int ny, int Nx, int Ny, const double s[3], const double d[3], double shz,
	std::vector<double>& Dd, std::vector<double>& Do)
{
	const int A = kAsymptoticDistance;
	const int dx = std::min(nx, A), dy = std::min(ny, A);
	const size_t NN = (size_t)Nx * Ny;
	Dd.assign(NN * 3, 0.0); Do.assign(NN * 3, 0.0);
	const bool regular = same_cell(s, d);
	// the regular path runs on the source cell (Boris/DemagTFunc_Irregular.cpp:11)
	const double hx = regular ? s[0] : d[0], hy = regular ? s[1] : d[1], hz = regular ? s[2] : d[2];
	auto lin = [](double h, double o) { return [h, o](int i) { return i * h + o; }; };

	Grid3 fxx, fyy, fzz, gxy, gxz, gyz, fxx_d, fyy_d, fzz_d, gxy_d, gxz_d, gyz_d;
	if (regular) {
		fxx = fill_grid_at(nw_f, dx, dy, 1, lin(hx, 0), lin(hy, 0), lin(hz, shz));
		fyy = fill_grid_at(nw_f, dy, dx, 1, lin(hy, 0), lin(hx, 0), lin(hz, shz));
		fzz = fill_grid_at(nw_f, 1, dy, dx, lin(hz, shz), lin(hy, 0), lin(hx, 0));
		gxy = fill_grid_at(nw_g, dx, dy, 1, lin(hx, 0), lin(hy, 0), lin(hz, shz));
		gxz = fill_grid_at(nw_g, dx, 1, dy, lin(hx, 0), lin(hz, shz), lin(hy, 0));
		gyz = fill_grid_at(nw_g, dy, 1, dx, lin(hy, 0), lin(hz, shz), lin(hx, 0));
	}
	else {
		const double del = s[2] - d[2];
		// plane positions -1, 0, +1 of each grid (xx, yy: -sz, 0, +dz; zz, g: -dz, 0, +sz)
		auto p_sd = [&](int k) { return (k < 0 ? -s[2] : k > 0 ? d[2] : 0.0) + shz; };
		auto p_ds = [&](int k) { return (k < 0 ? -d[2] : k > 0 ? s[2] : 0.0) + shz; };
		auto c_m = [&](int) { return -del + shz; };
		auto c_p = [&](int) { return del + shz; };
		fxx = fill_grid_at(nw_f, dx, dy, 1, lin(hx, 0), lin(hy, 0), p_sd);
		fyy = fill_grid_at(nw_f, dy, dx, 1, lin(hy, 0), lin(hx, 0), p_sd);
		fzz = fill_grid_at(nw_f, 1, dy, dx, p_ds, lin(hy, 0), lin(hx, 0));
		gxy = fill_grid_at(nw_g, dx, dy, 1, lin(hx, 0), lin(hy, 0), p_ds);
		gxz = fill_grid_at(nw_g, dx, 1, dy, lin(hx, 0), p_ds, lin(hy, 0));
		gyz = fill_grid_at(nw_g, dy, 1, dx, lin(hy, 0), p_ds, lin(hx, 0));
		// single extra planes (stored at plane index 0 of a one-plane grid)
		fxx_d = fill_grid_at(nw_f, dx, dy, -1, lin(hx, 0), lin(hy, 0), c_m);
		fyy_d = fill_grid_at(nw_f, dy, dx, -1, lin(hy, 0), lin(hx, 0), c_m);
		fzz_d = fill_grid_at(nw_f, -1, dy, dx, c_p, lin(hy, 0), lin(hx, 0));
		gxy_d = fill_grid_at(nw_g, dx, dy, -1, lin(hx, 0), lin(hy, 0), c_p);
		gxz_d = fill_grid_at(nw_g, dx, -1, dy, lin(hx, 0), c_p, lin(hy, 0));
		gyz_d = fill_grid_at(nw_g, dy, -1, dx, lin(hy, 0), c_p, lin(hx, 0));
	}
	AsymptDiag adx(hx, hy, hz), ady(hy, hx, hz), adz(hz, hy, hx);
	AsymptOffDiag aoxy(hx, hy, hz), aoxz(hx, hz, hy), aoyz(hy, hz, hx);
	const double tau = hx * hy * hz;
	const double far_scale = regular ? 1.0 : s[2];
	const double ks = std::fabs(shz / hz);
	static const int mir[4][2] = { {1,1},{-1,1},{1,-1},{-1,-1} };

#pragma omp parallel for schedule(dynamic, 1)
	for (int j = 0; j < ny; j++) {
		for (int i = 0; i < nx; i++) {
			double vd[3], vo[3];
			if (A > 0 && (i >= A || j >= A || floor_eps(ks) >= A || floor_eps(i * i + j * j + ks * ks) >= A * A)) {
				vd[0] = (-adx.eval(i * hx, j * hy, shz)) * far_scale;
				vd[1] = (-ady.eval(j * hy, i * hx, shz)) * far_scale;
				vd[2] = (-adz.eval(shz, j * hy, i * hx)) * far_scale;
				vo[0] = (-aoxy.eval(i * hx, j * hy, shz)) * far_scale;
				vo[1] = (-aoxz.eval(i * hx, shz, j * hy)) * far_scale;
				vo[2] = (-aoyz.eval(j * hy, shz, i * hx)) * far_scale;
			}
			else if (regular) {
				vd[0] = L27(fxx, i, j, 0, hx, hy, hz) * -1;
				vd[1] = L27(fyy, j, i, 0, hy, hx, hz) * -1;
				vd[2] = L27(fzz, 0, j, i, hz, hy, hx) * -1;
				vo[0] = L27(gxy, i, j, 0, hx, hy, hz) * -1;
				vo[1] = L27(gxz, i, 0, j, hx, hz, hy) * -1;
				vo[2] = L27(gyz, j, 0, i, hy, hz, hx) * -1;
			}
			else {
				vd[0] = L36_plane3(fxx, fxx_d, i, j, tau) * -1;
				vd[1] = L36_plane3(fyy, fyy_d, j, i, tau) * -1;
				vd[2] = L36_zz(fzz, fzz_d, j, i, tau) * -1;
				vo[0] = L36_plane3(gxy, gxy_d, i, j, tau) * -1;
				vo[1] = L36_plane2(gxz, gxz_d, i, j, tau) * -1;
				vo[2] = L36_plane2(gyz, gyz_d, j, i, tau) * -1;
			}
			for (int m = 0; m < 4; m++) {
				int last = m;
				for (int q = m + 1; q < 4; q++)
					if ((mir[q][0] * i == mir[m][0] * i) && (mir[q][1] * j == mir[m][1] * j)) last = q;
				if (last != m) continue;
				const int si = mir[m][0], sj = mir[m][1];
				const size_t idx = wrap(si * i, Nx) + wrap(sj * j, Ny) * Nx;
				for (int c = 0; c < 3; c++) Dd[idx * 3 + c] = vd[c];
				Do[idx * 3 + 0] = vo[0] * (si * sj);
				Do[idx * 3 + 1] = vo[1] * si;
				Do[idx * 3 + 2] = vo[2] * sj;
			}
		}
	}
}

// 2D tensor between source cells s and destination cells d (normalised, equal in x and y)
// for a destination layer shifted by (shx, shy, shz) with shx or shy nonzero: the general
// branch of Boris/DemagTFunc_Shifted.cpp:235-343 (regular, grids of
// Boris/DemagTFunc_fg.cpp:303-350) and Boris/DemagTFunc_Irregular.cpp:5-235 (irregular, grids
// :411-506), no symmetry used: every offset i in (-nx, nx), j in (-ny, ny) at position
// i hx + shx, j hy + shy, stored at (i mod Nx, j mod Ny), -N.  The near field uses the same
// 27- and 36-term stencils as the z-shifted tensor on grids over the offsets the near field
// reads (the reference fills them over all of [-n, n]), the far field the series
// at the shifted position (scaled by s.z when irregular).
inline void tensor2d_shifted(int nx, int ny, int Nx, int Ny, const double s[3], const double d[3], const double sh[3],
	std::vector<double>& Dd, std::vector<double>& Do)
{
	const int A = kAsymptoticDistance;
	const size_t NN = (size_t)Nx * Ny;
	Dd.assign(NN * 3, 0.0); Do.assign(NN * 3, 0.0);
	const bool regular = same_cell(s, d);
	const double hx = regular ? s[0] : d[0], hy = regular ? s[1] : d[1], hz = regular ? s[2] : d[2];
	const double shx = sh[0], shy = sh[1], shz = sh[2];
	// grid index q of offset i: q = i - lo - 1 over [lo, hi], the offsets the near field can
	// read (|i + sh/h| below the far-field distance, +-1 for the stencil) within [-n, n]
	auto lo_of = [&](int n, double f) { return std::max(-n, (int)std::floor(-A - f) - 2); };
	auto hi_of = [&](int n, double f) { return std::min(n, (int)std::ceil(A - f) + 2); };
	const int lx = lo_of(nx, shx / hx), hxi = hi_of(nx, shx / hx), ly = lo_of(ny, shy / hy), hyi = hi_of(ny, shy / hy);
	const int dx = hxi - lx - 1, dy = hyi - ly - 1;
	auto ax = [&](double h, double o, int lo) { return [h, o, lo](int q) { return (q + lo + 1) * h + o; }; };
	auto lin = [](double h, double o) { return [h, o](int k) { return k * h + o; }; };
	const auto X = ax(hx, shx, lx), Y = ax(hy, shy, ly);
	Grid3 fxx, fyy, fzz, gxy, gxz, gyz, fxx_d, fyy_d, fzz_d, gxy_d, gxz_d, gyz_d;
	if (regular) {
		fxx = fill_grid_at(nw_f, dx, dy, 1, X, Y, lin(hz, shz));
		fyy = fill_grid_at(nw_f, dy, dx, 1, Y, X, lin(hz, shz));
		fzz = fill_grid_at(nw_f, 1, dy, dx, lin(hz, shz), Y, X);
		gxy = fill_grid_at(nw_g, dx, dy, 1, X, Y, lin(hz, shz));
		gxz = fill_grid_at(nw_g, dx, 1, dy, X, lin(hz, shz), Y);
		gyz = fill_grid_at(nw_g, dy, 1, dx, Y, lin(hz, shz), X);
	}
	else {
		const double del = s[2] - d[2];
		auto p_sd = [&](int k) { return (k < 0 ? -s[2] : k > 0 ? d[2] : 0.0) + shz; };
		auto p_ds = [&](int k) { return (k < 0 ? -d[2] : k > 0 ? s[2] : 0.0) + shz; };
		auto c_m = [&](int) { return -del + shz; };
		auto c_p = [&](int) { return del + shz; };
		fxx = fill_grid_at(nw_f, dx, dy, 1, X, Y, p_sd);
		fyy = fill_grid_at(nw_f, dy, dx, 1, Y, X, p_sd);
		fzz = fill_grid_at(nw_f, 1, dy, dx, p_ds, Y, X);
		gxy = fill_grid_at(nw_g, dx, dy, 1, X, Y, p_ds);
		gxz = fill_grid_at(nw_g, dx, 1, dy, X, p_ds, Y);
		gyz = fill_grid_at(nw_g, dy, 1, dx, Y, p_ds, X);
		fxx_d = fill_grid_at(nw_f, dx, dy, -1, X, Y, c_m);
		fyy_d = fill_grid_at(nw_f, dy, dx, -1, Y, X, c_m);
		fzz_d = fill_grid_at(nw_f, -1, dy, dx, c_p, Y, X);
		gxy_d = fill_grid_at(nw_g, dx, dy, -1, X, Y, c_p);
		gxz_d = fill_grid_at(nw_g, dx, -1, dy, X, c_p, Y);
		gyz_d = fill_grid_at(nw_g, dy, -1, dx, Y, c_p, X);
	}
	AsymptDiag adx(hx, hy, hz), ady(hy, hx, hz), adz(hz, hy, hx);
	AsymptOffDiag aoxy(hx, hy, hz), aoxz(hx, hz, hy), aoyz(hy, hz, hx);
	const double tau = hx * hy * hz;
	const double far_scale = regular ? 1.0 : s[2];
	const double ks = std::fabs(shz / hz);

#pragma omp parallel for schedule(dynamic, 1)
	for (int j = -ny + 1; j < ny; j++) {
		for (int i = -nx + 1; i < nx; i++) {
			const double is = std::fabs(i + shx / hx), js = std::fabs(j + shy / hy);
			const double x = i * hx + shx, y = j * hy + shy;
			double vd[3], vo[3];
			if (A > 0 && (floor_eps(is) >= A || floor_eps(js) >= A || floor_eps(ks) >= A || floor_eps(is * is + js * js + ks * ks) >= A * A)) {
				vd[0] = (-adx.eval(x, y, shz)) * far_scale;
				vd[1] = (-ady.eval(y, x, shz)) * far_scale;
				vd[2] = (-adz.eval(shz, y, x)) * far_scale;
				vo[0] = (-aoxy.eval(x, y, shz)) * far_scale;
				vo[1] = (-aoxz.eval(x, shz, y)) * far_scale;
				vo[2] = (-aoyz.eval(y, shz, x)) * far_scale;
			}
			else {
				const int qi = i - lx - 1, qj = j - ly - 1;
				if (regular) {
					vd[0] = L27(fxx, qi, qj, 0, hx, hy, hz) * -1;
					vd[1] = L27(fyy, qj, qi, 0, hy, hx, hz) * -1;
					vd[2] = L27(fzz, 0, qj, qi, hz, hy, hx) * -1;
					vo[0] = L27(gxy, qi, qj, 0, hx, hy, hz) * -1;
					vo[1] = L27(gxz, qi, 0, qj, hx, hz, hy) * -1;
					vo[2] = L27(gyz, qj, 0, qi, hy, hz, hx) * -1;
				}
				else {
					vd[0] = L36_plane3(fxx, fxx_d, qi, qj, tau) * -1;
					vd[1] = L36_plane3(fyy, fyy_d, qj, qi, tau) * -1;
					vd[2] = L36_zz(fzz, fzz_d, qj, qi, tau) * -1;
					vo[0] = L36_plane3(gxy, gxy_d, qi, qj, tau) * -1;
					vo[1] = L36_plane2(gxz, gxz_d, qi, qj, tau) * -1;
					vo[2] = L36_plane2(gyz, gyz_d, qj, qi, tau) * -1;
				}
			}
			const size_t idx = wrap(i, Nx) + wrap(j, Ny) * Nx;
			for (int c = 0; c < 3; c++) { Dd[idx * 3 + c] = vd[c]; Do[idx * 3 + c] = vo[c]; }
		}
	}
}

// Complex kernels of a general shifted 2D tensor (Boris/DemagKernelCollectionCUDA_Calc.cpp:
// 571-745, the Complex_Full kernels): x real-input FFT keeping kx in [0, Nx/2], then y complex
// FFT, all six components complex; output Nc x Ny x 6 complex (kx fastest), components xx,
// yy, zz, xy, xz, yz.
inline void kernels2d_complex(int Nx, int Ny, const std::vector<double>& Dd, const std::vector<double>& Do,
	std::vector<std::complex<double>>& K)
{
	typedef std::complex<double> C;
	const int Nc = Nx / 2 + 1;
	K.assign((size_t)Nc * Ny * 6, C(0, 0));
	Fft<double> fx(Nx), fy(Ny);
	std::vector<C> a(Nx), sc(Nx), b(Ny), sb(Ny);
	for (int c = 0; c < 6; c++) {
		const std::vector<double>& D = c < 3 ? Dd : Do;
		for (int j = 0; j < Ny; j++) {
			for (int i = 0; i < Nx; i++) a[i] = C(D[((size_t)i + (size_t)j * Nx) * 3 + c % 3], 0.0);
			fx.run(a.data(), sc.data(), -1);
			for (int i = 0; i < Nc; i++) K[((size_t)i + (size_t)j * Nc) * 6 + c] = a[i];
		}
		for (int i = 0; i < Nc; i++) {
			for (int j = 0; j < Ny; j++) b[j] = K[((size_t)i + (size_t)j * Nc) * 6 + c];
			fy.run(b.data(), sb.data(), -1);
			for (int j = 0; j < Ny; j++) K[((size_t)i + (size_t)j * Nc) * 6 + c] = b[j];
		}
	}
}

// The same tensor with periodic images along x (px) and y (py), no images along z: the
// z-shift branch of Boris/DemagTFunc_Shifted_PBC.cpp:317-460 / 464-600 (regular) and
// Boris/DemagTFunc_Irregular_PBC.cpp:5-120 (irregular).  Along a periodic axis N = n and
// every offset i0 in [0, N) sums the images i = i_img N + i0, i_img = -p..p (i_img outer, j_img
// inner), each at |i| by the near-field stencil (grids out to the far-field distance) or the
// far-field series, the off-diagonal terms signed by sign(i) sign(j) (xy), sign(i) (xz),
// sign(j) (yz); along an open axis only [0, N/2) is computed and mirrored as without images.
inline void tensor2d_zshift_pbc(int Nx, int Ny, const double s[3], const double d[3], double shz, int px, int py,
	std::vector<double>& Dd, std::vector<double>& Do)
{
	const int A = kAsymptoticDistance;
	const int ex = px ? Nx : Nx / 2, ey = py ? Ny : Ny / 2;   // computed offsets per axis
	const int dx = px ? A : std::min(ex, A), dy = py ? A : std::min(ey, A);
	const size_t NN = (size_t)Nx * Ny;
	Dd.assign(NN * 3, 0.0); Do.assign(NN * 3, 0.0);
	const bool regular = same_cell(s, d);
	const double hx = regular ? s[0] : d[0], hy = regular ? s[1] : d[1], hz = regular ? s[2] : d[2];
	auto lin = [](double h, double o) { return [h, o](int i) { return i * h + o; }; };
	Grid3 fxx, fyy, fzz, gxy, gxz, gyz, fxx_d, fyy_d, fzz_d, gxy_d, gxz_d, gyz_d;
	if (regular) {
		fxx = fill_grid_at(nw_f, dx, dy, 1, lin(hx, 0), lin(hy, 0), lin(hz, shz));
		fyy = fill_grid_at(nw_f, dy, dx, 1, lin(hy, 0), lin(hx, 0), lin(hz, shz));
		fzz = fill_grid_at(nw_f, 1, dy, dx, lin(hz, shz), lin(hy, 0), lin(hx, 0));
		gxy = fill_grid_at(nw_g, dx, dy, 1, lin(hx, 0), lin(hy, 0), lin(hz, shz));
		gxz = fill_grid_at(nw_g, dx, 1, dy, lin(hx, 0), lin(hz, shz), lin(hy, 0));
		gyz = fill_grid_at(nw_g, dy, 1, dx, lin(hy, 0), lin(hz, shz), lin(hx, 0));
	}
	else {
		const double del = s[2] - d[2];
		auto p_sd = [&](int k) { return (k < 0 ? -s[2] : k > 0 ? d[2] : 0.0) + shz; };
		auto p_ds = [&](int k) { return (k < 0 ? -d[2] : k > 0 ? s[2] : 0.0) + shz; };
		auto c_m = [&](int) { return -del + shz; };
		auto c_p = [&](int) { return del + shz; };
		fxx = fill_grid_at(nw_f, dx, dy, 1, lin(hx, 0), lin(hy, 0), p_sd);
		fyy = fill_grid_at(nw_f, dy, dx, 1, lin(hy, 0), lin(hx, 0), p_sd);
		fzz = fill_grid_at(nw_f, 1, dy, dx, p_ds, lin(hy, 0), lin(hx, 0));
		gxy = fill_grid_at(nw_g, dx, dy, 1, lin(hx, 0), lin(hy, 0), p_ds);
		gxz = fill_grid_at(nw_g, dx, 1, dy, lin(hx, 0), p_ds, lin(hy, 0));
		gyz = fill_grid_at(nw_g, dy, 1, dx, lin(hy, 0), p_ds, lin(hx, 0));
		fxx_d = fill_grid_at(nw_f, dx, dy, -1, lin(hx, 0), lin(hy, 0), c_m);
		fyy_d = fill_grid_at(nw_f, dy, dx, -1, lin(hy, 0), lin(hx, 0), c_m);
		fzz_d = fill_grid_at(nw_f, -1, dy, dx, c_p, lin(hy, 0), lin(hx, 0));
		gxy_d = fill_grid_at(nw_g, dx, dy, -1, lin(hx, 0), lin(hy, 0), c_p);
		gxz_d = fill_grid_at(nw_g, dx, -1, dy, lin(hx, 0), c_p, lin(hy, 0));
		gyz_d = fill_grid_at(nw_g, dy, -1, dx, lin(hy, 0), c_p, lin(hx, 0));
	}
	AsymptDiag adx(hx, hy, hz), ady(hy, hx, hz), adz(hz, hy, hx);
	AsymptOffDiag aoxy(hx, hy, hz), aoxz(hx, hz, hy), aoyz(hy, hz, hx);
	const double tau = hx * hy * hz;
	const double far_scale = regular ? 1.0 : s[2];
	const double ks = std::fabs(shz / hz);

#pragma omp parallel for schedule(dynamic, 1)
	for (int j0 = 0; j0 < ey; j0++) {
		for (int i0 = 0; i0 < ex; i0++) {
			double vd[3] = { 0, 0, 0 }, vo[3] = { 0, 0, 0 };
			for (int ii = -px; ii <= px; ii++)
				for (int jj = -py; jj <= py; jj++) {
					const int is = ii * Nx + i0, js = jj * Ny + j0;
					const double si = (is > 0) - (is < 0), sj = (js > 0) - (js < 0);
					const int i = std::abs(is), j = std::abs(js);
					double td[3], to[3];
					if (A > 0 && (i >= A || j >= A || floor_eps(ks) >= A || floor_eps(i * i + j * j + ks * ks) >= A * A)) {
						td[0] = (-adx.eval(i * hx, j * hy, shz)) * far_scale;
						td[1] = (-ady.eval(j * hy, i * hx, shz)) * far_scale;
						td[2] = (-adz.eval(shz, j * hy, i * hx)) * far_scale;
						to[0] = (-aoxy.eval(i * hx, j * hy, shz)) * far_scale;
						to[1] = (-aoxz.eval(i * hx, shz, j * hy)) * far_scale;
						to[2] = (-aoyz.eval(j * hy, shz, i * hx)) * far_scale;
					}
					else if (regular) {
						td[0] = L27(fxx, i, j, 0, hx, hy, hz) * -1;
						td[1] = L27(fyy, j, i, 0, hy, hx, hz) * -1;
						td[2] = L27(fzz, 0, j, i, hz, hy, hx) * -1;
						to[0] = L27(gxy, i, j, 0, hx, hy, hz) * -1;
						to[1] = L27(gxz, i, 0, j, hx, hz, hy) * -1;
						to[2] = L27(gyz, j, 0, i, hy, hz, hx) * -1;
					}
					else {
						td[0] = L36_plane3(fxx, fxx_d, i, j, tau) * -1;
						td[1] = L36_plane3(fyy, fyy_d, j, i, tau) * -1;
						td[2] = L36_zz(fzz, fzz_d, j, i, tau) * -1;
						to[0] = L36_plane3(gxy, gxy_d, i, j, tau) * -1;
						to[1] = L36_plane2(gxz, gxz_d, i, j, tau) * -1;
						to[2] = L36_plane2(gyz, gyz_d, j, i, tau) * -1;
					}
					for (int c = 0; c < 3; c++) vd[c] += td[c];
					vo[0] += to[0] * (si * sj); vo[1] += to[1] * si; vo[2] += to[2] * sj;
				}
			const size_t idx = (size_t)i0 + (size_t)j0 * Nx;
			auto put = [&](size_t q, double ox, double oy, double oz) {
				for (int c = 0; c < 3; c++) Dd[q * 3 + c] = vd[c];
				Do[q * 3 + 0] = vo[0] * ox; Do[q * 3 + 1] = vo[1] * oy; Do[q * 3 + 2] = vo[2] * oz;
			};
			put(idx, 1, 1, 1);
			const size_t mx = (size_t)((Nx - i0) % Nx), my = (size_t)((Ny - j0) % Ny);
			if (!px) {
				if (i0) put(mx + (size_t)j0 * Nx, -1, -1, 1);
				if (!py && i0 && j0) put(mx + my * Nx, 1, -1, -1);
			}
			if (!py && j0) put((size_t)i0 + my * Nx, -1, 1, -1);
		}
	}
}

// Real kernels of a z-shifted 2D tensor (Boris/DemagKernelCollection_Calc.cpp:288-420):
// x then y real-input FFTs; diagonal parts Re, Re; off-diagonal (xy, xz, yz) take
// (Im, Im, Re) along x and (-Im, Re, Im) along y.  The stored xz and yz values are the
// coefficients of i (the product multiplies them by i).  Output (Ny/2+1) x (Nx/2+1) x 3
// each, kx fastest.
inline void kernels2d_zshift(int Nx, int Ny, const std::vector<double>& Dd, const std::vector<double>& Do,
	std::vector<double>& Kd, std::vector<double>& Ko)
{
	typedef std::complex<double> C;
	const int Nc = Nx / 2 + 1, Nyh = Ny / 2 + 1;
	std::vector<double> Td((size_t)Nc * Ny * 3), To((size_t)Nc * Ny * 3);
	Fft<double> fx(Nx), fy(Ny);
	{
		std::vector<C> a(Nx), sc(Nx);
		for (int j = 0; j < Ny; j++)
			for (int c = 0; c < 6; c++) {
				const std::vector<double>& D = c < 3 ? Dd : Do;
				for (int i = 0; i < Nx; i++) a[i] = C(D[((size_t)i + (size_t)j * Nx) * 3 + c % 3], 0.0);
				fx.run(a.data(), sc.data(), -1);
				for (int i = 0; i < Nc; i++) {
					double v = (c < 3 || c == 5) ? a[i].real() : a[i].imag();
					(c < 3 ? Td : To)[((size_t)i + (size_t)j * Nc) * 3 + c % 3] = v;
				}
			}
	}
	Kd.assign((size_t)Nc * Nyh * 3, 0.0); Ko.assign((size_t)Nc * Nyh * 3, 0.0);
	{
		std::vector<C> a(Ny), sc(Ny);
		for (int i = 0; i < Nc; i++)
			for (int c = 0; c < 6; c++) {
				const std::vector<double>& T = c < 3 ? Td : To;
				for (int j = 0; j < Ny; j++) a[j] = C(T[((size_t)i + (size_t)j * Nc) * 3 + c % 3], 0.0);
				fy.run(a.data(), sc.data(), -1);
				for (int j = 0; j < Nyh; j++) {
					double v = c < 3 ? a[j].real() : c == 3 ? -a[j].imag() : c == 4 ? a[j].real() : a[j].imag();
					(c < 3 ? Kd : Ko)[((size_t)i + (size_t)j * Nc) * 3 + c % 3] = v;
				}
			}
	}
}

// Multilayered 2D demag (SDemag multilayered convolution with every layer a 2D mesh of
// the same nx x ny cells, layers stacked along z): H_d = sum_s K_ds * M_s
// (Boris/SDemag.cpp:173-300 for the layer rectangles, Boris/DemagKernelCollection_Calc.cpp
// for the kernels, Boris/DemagKernelCollection_Mult.cpp:12-500 for the product; the
// inverse-shift kernel reuse of the reference is not needed here: every pair's kernel is
// computed directly).  Layer l has cells (hx, hy, hz[l]) and its bottom face at z0[l].
template <typename R>
struct MLayer {
	typedef std::complex<R> C;
	int L, nx, ny, Nx, Ny, Nc, Nyh;
	// per (d, s) pair: self (d == s) -> Kd + Kxy (2D self kernel); else Kd + Ko (z-shifted),
	// or Kc (complex, Nc x Ny x 6) for a pair whose layers are offset in-plane
	std::vector<std::vector<R>> Kd, Ko;
	std::vector<std::vector<C>> Kc;
	std::vector<int> self;
	Fft<R> fx, fy;

	// px, py: periodic images along x and y (SDemag::Set_PBC, Boris/SDemag.cpp:403-408): N = n
	// along a periodic axis, the self kernels are the periodic 2D kernels and the pair
	// kernels the periodic z-shifted ones (Boris/DemagKernelCollection_Calc.cpp with pbc_images)
	// rx, ry (optional, m): in-plane offsets of the layers' grids, so a pair (d, s) with
	// rx[d] != rx[s] or ry[d] != ry[s] couples through the tensor shifted by the difference
	// (the reference's Complex_Full kernels, Boris/DemagKernelCollectionCUDA_Calc.cpp:571-745)
	MLayer(int L_, int nx_, int ny_, double hx, double hy, const double* hz, const double* z0, int px = 0, int py = 0,
		const double* rx = nullptr, const double* ry = nullptr)
		: L(L_), nx(nx_), ny(ny_)
	{
		Nx = px ? nx : 2 * nx; Ny = py ? ny : 2 * ny; Nc = Nx / 2 + 1; Nyh = Ny / 2 + 1;
		const int pbc3[3] = { px, py, 0 };
		const bool periodic = px || py;
		fx = Fft<R>(Nx); fy = Fft<R>(Ny);
		double hmax = 0;
		for (int l = 0; l < L; l++) hmax = std::max(hmax, std::max(hx, std::max(hy, hz[l])));
		Kd.resize((size_t)L * L); Ko.resize((size_t)L * L); Kc.resize((size_t)L * L); self.assign((size_t)L * L, 0);
		for (int d = 0; d < L; d++)
			for (int s = 0; s < L; s++) {
				const size_t p = (size_t)d * L + s;
				if (d == s) {
					Kernels K = make_kernels(nx, ny, 1, Nx, Ny, 1, hx, hy, hz[d], periodic ? pbc3 : nullptr);
					Kd[p].assign(K.Kdiag.begin(), K.Kdiag.end());
					Ko[p].assign(K.Kodiag.begin(), K.Kodiag.end());
					self[p] = 1;
				}
				else {
					const double sc[3] = { hx / hmax, hy / hmax, hz[s] / hmax }, dc[3] = { hx / hmax, hy / hmax, hz[d] / hmax };
					std::vector<double> Dd, Do, kd, ko;
					const double sx = rx ? (rx[d] - rx[s]) / hmax : 0.0, sy = ry ? (ry[d] - ry[s]) / hmax : 0.0;
					if (std::fabs(sx) > 1e-12 || std::fabs(sy) > 1e-12) {
						if (periodic) throw std::runtime_error("in-plane offset layers with periodic images");
						const double sh[3] = { sx, sy, (z0[d] - z0[s]) / hmax };
						tensor2d_shifted(nx, ny, Nx, Ny, sc, dc, sh, Dd, Do);
						std::vector<std::complex<double>> kc;
						kernels2d_complex(Nx, Ny, Dd, Do, kc);
						Kc[p].assign(kc.begin(), kc.end());
						continue;
					}
					if (periodic) tensor2d_zshift_pbc(Nx, Ny, sc, dc, (z0[d] - z0[s]) / hmax, px, py, Dd, Do);
					else tensor2d_zshift(nx, ny, Nx, Ny, sc, dc, (z0[d] - z0[s]) / hmax, Dd, Do);
					kernels2d_zshift(Nx, Ny, Dd, Do, kd, ko);
					Kd[p].assign(kd.begin(), kd.end());
					Ko[p].assign(ko.begin(), ko.end());
				}
			}
	}

	// M, H: L layers of nx*ny cells of 3 reals (layer-major, x fastest).  H = H_demag, or with
	// transpose the transposed operator (N^T)_sd = N_ds^T: in Fourier space the pair (d, s)
	// kernel K_ds serves output s from input d with its imaginary xz, yz parts negated
	// (the real-space kernel N_ds(-r)^T).
	void convolute(const R* M, R* H, bool transpose = false)
	{
		const size_t nl = (size_t)nx * ny, fl = (size_t)Nc * Ny;
		std::vector<C> F((size_t)L * 3 * fl), O((size_t)3 * fl);
		auto fi = [&](int l, int c, int i, int j) { return ((size_t)l * 3 + c) * fl + (size_t)i + (size_t)j * Nc; };
		// forward: x R2C then y for each layer and component
		for (int l = 0; l < L; l++)
			for (int c = 0; c < 3; c++) {
				std::vector<C> a(Nx), sx(Nx), b(Ny), sy(Ny);
				for (int j = 0; j < ny; j++) {
					for (int i = 0; i < Nx; i++) a[i] = C(i < nx ? M[(l * nl + (size_t)i + (size_t)j * nx) * 3 + c] : (R)0, (R)0);
					fx.run(a.data(), sx.data(), -1);
					for (int i = 0; i < Nc; i++) F[fi(l, c, i, j)] = a[i];
				}
				for (int i = 0; i < Nc; i++) {
					for (int j = 0; j < Ny; j++) b[j] = j < ny ? F[fi(l, c, i, j)] : C(0, 0);
					fy.run(b.data(), sy.data(), -1);
					for (int j = 0; j < Ny; j++) F[fi(l, c, i, j)] = b[j];
				}
			}
		const C I1(0, 1);
		for (int d = 0; d < L; d++) {
			std::fill(O.begin(), O.end(), C(0, 0));
			// self contribution first, then the others in layer order
			for (int q = 0; q < L; q++) {
				const int s = q == 0 ? d : (q <= d ? q - 1 : q);
				const size_t p = transpose ? (size_t)s * L + d : (size_t)d * L + s;
				const R tz = transpose ? (R)-1 : (R)1;
				for (int j = 0; j < Ny; j++) {
					const int jj = j <= Ny / 2 ? j : Ny - j;
					const R sy = j <= Ny / 2 ? (R)1 : (R)-1;
					for (int i = 0; i < Nc; i++) {
						const size_t k = (size_t)i + (size_t)jj * Nc, o = (size_t)i + (size_t)j * Nc;
						const C fxv = F[fi(s, 0, i, j)], fyv = F[fi(s, 1, i, j)], fzv = F[fi(s, 2, i, j)];
						if (!Kc[p].empty()) {
							// complex kernel (transpose: its conjugate)
							const size_t kk = ((size_t)i + (size_t)j * Nc) * 6;
							auto K = [&](int c) { return transpose ? std::conj(Kc[p][kk + c]) : Kc[p][kk + c]; };
							O[o] += K(0) * fxv + K(3) * fyv + K(4) * fzv;
							O[fl + o] += K(3) * fxv + K(1) * fyv + K(5) * fzv;
							O[2 * fl + o] += K(4) * fxv + K(5) * fyv + K(2) * fzv;
							continue;
						}
						const R kxx = Kd[p][k * 3], kyy = Kd[p][k * 3 + 1], kzz = Kd[p][k * 3 + 2];
						if (self[p]) {
							const R kxy = sy * Ko[p][k];
							O[o] += kxx * fxv + kxy * fyv;
							O[fl + o] += kxy * fxv + kyy * fyv;
							O[2 * fl + o] += kzz * fzv;
						}
						else {
							const R kxy = sy * Ko[p][k * 3], kxz = tz * Ko[p][k * 3 + 1], kyz = tz * sy * Ko[p][k * 3 + 2];
							O[o] += kxx * fxv + kxy * fyv + I1 * (kxz * fzv);
							O[fl + o] += kxy * fxv + kyy * fyv + I1 * (kyz * fzv);
							O[2 * fl + o] += I1 * (kxz * fxv) + I1 * (kyz * fyv) + kzz * fzv;
						}
					}
				}
			}
			// inverse y then x C2R, scale
			const R scale = (R)1 / (R)((double)Nx * Ny);
			for (int c = 0; c < 3; c++) {
				std::vector<C> a(Nx), sx(Nx), b(Ny), sy(Ny);
				for (int i = 0; i < Nc; i++) {
					for (int j = 0; j < Ny; j++) b[j] = O[c * fl + (size_t)i + (size_t)j * Nc];
					fy.run(b.data(), sy.data(), +1);
					for (int j = 0; j < ny; j++) O[c * fl + (size_t)i + (size_t)j * Nc] = b[j];
				}
				for (int j = 0; j < ny; j++) {
					for (int i = 0; i < Nc; i++) a[i] = O[c * fl + (size_t)i + (size_t)j * Nc];
					for (int i = Nc; i < Nx; i++) a[i] = std::conj(a[Nx - i]);
					fx.run(a.data(), sx.data(), +1);
					for (int i = 0; i < nx; i++) H[(d * nl + (size_t)i + (size_t)j * nx) * 3 + c] = a[i].real() * scale;
				}
			}
		}
	}
};

}   // namespace oracle
