// ORACLE — TEST INFRASTRUCTURE ONLY.
// CPU FP64 restatement of the reference's Newell demag-tensor setup.  Only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg may load or call anything
// under oracle/; the product (boris2_amd/) never does.
//
// Parity status of each function (pins are in tests/test_oracle_tensor.py):
//   nw_f, nw_g, ldia, lodia, asympt_*, tensor3d/tensor2d, self_demag:
//       pinned bit-for-bit (and to 1e-15 rel) against oracle/_ref (the reference's own
//       DemagTFunc*.cpp compiled here) through committed golden vectors; also against
//       closed forms (cube self-demag -1/3, trace -1, far-field dipole).
//   tensor3d_pbc/tensor2d_pbc (periodic images): pinned bit-for-bit against oracle/_ref
//       (the reference's DemagTFunc_PBC.cpp) through tests/golden/pbc_golden.npz, and by
//       the trace identity (tests/test_oracle_pbc.py).
#pragma once

#include <cmath>
#include <cstdint>
#include <vector>
#include <algorithm>

namespace oracle {

static const double kPI = 3.1415926535897932384626433833;   // BorisLib/Funcs_Math_base.h:8
static const int kAsymptoticDistance = 40;                    // Boris/DemagTFunc_Defs.h:9
// BorisLib/Funcs_Aux_base.h:9 defines the zero threshold as the float literal 1e-12F.
static const double kEps = (double)1e-12F;

inline bool nz(double v) { return v > kEps || v < -kEps; }

// Kahan summation with Neumaier's modification, in the order given
// (BorisLib/Funcs_Math.h:294-325).
inline double kn_sum(const double* v, int n)
{
	double s = 0.0, c = 0.0;
	for (int i = 0; i < n; i++) {
		double t = s + v[i];
		if (std::fabs(s) > std::fabs(v[i])) c += (s - t) + v[i];
		else c += (v[i] - t) + s;
		s = t;
	}
	return s + c;
}

// Newell f basis function (Boris/DemagTFunc_fg.cpp:10-43).
inline double nw_f(double x, double y, double z)
{
	double xx = x * x, yy = y * y, zz = z * z;
	double rxy = xx + yy, rxz = xx + zz;
	double R = std::sqrt(xx + yy + zz);
	double t[4];
	t[0] = (2.0 * xx - yy - zz) * R / 6.0;
	t[1] = 0.0;
	if (nz(rxz)) { double a = 2.0 * y * (y + R) / rxz; if (a > -1.0) t[1] = y * (zz - xx) * std::log1p(a) / 4.0; }
	t[2] = 0.0;
	if (nz(rxy)) { double a = 2.0 * z * (z + R) / rxy; if (a > -1.0) t[2] = z * (yy - xx) * std::log1p(a) / 4.0; }
	t[3] = nz(x) ? -x * y * z * std::atan(y * z / (x * R)) : 0.0;
	return kn_sum(t, 4);
}

// Newell g basis function (Boris/DemagTFunc_fg.cpp:45-92).
inline double nw_g(double x, double y, double z)
{
	double xx = x * x, yy = y * y, zz = z * z;
	double rxy = xx + yy, rxz = xx + zz, ryz = yy + zz;
	double R = std::sqrt(xx + yy + zz);
	double t[7];
	t[0] = -x * y * R / 3.0;
	t[1] = 0.0;
	if (nz(rxy)) { double a = 2.0 * z * (z + R) / rxy; if (a > -1.0) t[1] = x * y * z * std::log1p(a) / 2.0; }
	t[2] = 0.0;
	if (nz(ryz)) { double a = 2.0 * x * (x + R) / ryz; if (a > -1.0) t[2] = y * (3.0 * zz - yy) * std::log1p(a) / 12.0; }
	t[3] = 0.0;
	if (nz(rxz)) { double a = 2.0 * y * (y + R) / rxz; if (a > -1.0) t[3] = x * (3.0 * zz - xx) * std::log1p(a) / 12.0; }
	t[4] = nz(z) ? -zz * z * std::atan(x * y / (z * R)) / 6.0 : 0.0;
	t[5] = nz(y) ? -yy * z * std::atan(x * z / (y * R)) / 2.0 : 0.0;
	t[6] = nz(x) ? -xx * z * std::atan(y * z / (x * R)) / 2.0 : 0.0;
	return kn_sum(t, 7);
}

// Values of a basis function on the grid of offsets (-1..d) in each axis, stored with
// offset +1, x fastest (the reference's f_vals/g_vals, Boris/DemagTFunc_fg.cpp:94-260).
struct Grid3 {
	int nx = 0, ny = 0, nz_ = 0;
	std::vector<double> v;
	double at(int i, int j, int k) const { return v[(size_t)i + (size_t)j * nx + (size_t)k * nx * ny]; }
};

template <typename F>
inline Grid3 fill_grid(F fn, int dx, int dy, int dz, double hx, double hy, double hz)
{
	Grid3 g; g.nx = dx + 2; g.ny = dy + 2; g.nz_ = dz + 2;
	g.v.assign((size_t)g.nx * g.ny * g.nz_, 0.0);
#pragma omp parallel for collapse(2) schedule(static)
	for (int k = -1; k <= dz; k++)
		for (int j = -1; j <= dy; j++)
			for (int i = -1; i <= dx; i++)
				g.v[(size_t)(i + 1) + (size_t)(j + 1) * g.nx + (size_t)(k + 1) * g.nx * g.ny] = fn(i * hx, j * hy, k * hz);
	return g;
}

// 27-point second difference of a basis-function grid divided by 4*pi*hx*hy*hz.  This is
// both Ldia (on f) and Lodia (on g): Boris/DemagTFunc_L.cpp:8-57 and :59-110 use the same
// stencil and term order.
inline double L27(const Grid3& g, int i, int j, int k, double hx, double hy, double hz)
{
	i++; j++; k++;
	double t[27];
	t[0] = +8 * g.at(i, j, k);
	t[1] = -4 * g.at(i + 1, j, k);     t[2] = -4 * g.at(i - 1, j, k);
	t[3] = -4 * g.at(i, j + 1, k);     t[4] = -4 * g.at(i, j - 1, k);
	t[5] = -4 * g.at(i, j, k + 1);     t[6] = -4 * g.at(i, j, k - 1);
	t[7] = +2 * g.at(i - 1, j - 1, k); t[8] = +2 * g.at(i - 1, j + 1, k);
	t[9] = +2 * g.at(i + 1, j - 1, k); t[10] = +2 * g.at(i + 1, j + 1, k);
	t[11] = +2 * g.at(i - 1, j, k - 1); t[12] = +2 * g.at(i - 1, j, k + 1);
	t[13] = +2 * g.at(i + 1, j, k - 1); t[14] = +2 * g.at(i + 1, j, k + 1);
	t[15] = +2 * g.at(i, j - 1, k - 1); t[16] = +2 * g.at(i, j - 1, k + 1);
	t[17] = +2 * g.at(i, j + 1, k - 1); t[18] = +2 * g.at(i, j + 1, k + 1);
	t[19] = -1 * g.at(i - 1, j - 1, k - 1); t[20] = -1 * g.at(i - 1, j - 1, k + 1);
	t[21] = -1 * g.at(i - 1, j + 1, k - 1); t[22] = -1 * g.at(i + 1, j - 1, k - 1);
	t[23] = -1 * g.at(i - 1, j + 1, k + 1); t[24] = -1 * g.at(i + 1, j - 1, k + 1);
	t[25] = -1 * g.at(i + 1, j + 1, k - 1); t[26] = -1 * g.at(i + 1, j + 1, k + 1);
	return kn_sum(t, 27) / (4 * kPI * hx * hy * hz);
}

// Self-demag coefficient by direct evaluation of f (Boris/DemagTFunc_Self.cpp:6-44).
inline double self_demag_1(double hx, double hy, double hz)
{
	double t[27];
	t[0] = +8 * nw_f(0, 0, 0);
	t[1] = -4 * nw_f(hx, 0, 0);  t[2] = -4 * nw_f(-hx, 0, 0);
	t[3] = -4 * nw_f(0, hy, 0);  t[4] = -4 * nw_f(0, -hy, 0);
	t[5] = -4 * nw_f(0, 0, hz);  t[6] = -4 * nw_f(0, 0, -hz);
	t[7] = +2 * nw_f(-hx, -hy, 0); t[8] = +2 * nw_f(-hx, hy, 0);
	t[9] = +2 * nw_f(hx, -hy, 0);  t[10] = +2 * nw_f(hx, hy, 0);
	t[11] = +2 * nw_f(-hx, 0, -hz); t[12] = +2 * nw_f(-hx, 0, hz);
	t[13] = +2 * nw_f(hx, 0, -hz);  t[14] = +2 * nw_f(hx, 0, hz);
	t[15] = +2 * nw_f(0, -hy, -hz); t[16] = +2 * nw_f(0, -hy, hz);
	t[17] = +2 * nw_f(0, hy, -hz);  t[18] = +2 * nw_f(0, hy, hz);
	t[19] = -1 * nw_f(-hx, -hy, -hz); t[20] = -1 * nw_f(-hx, -hy, hz);
	t[21] = -1 * nw_f(-hx, hy, -hz);  t[22] = -1 * nw_f(hx, -hy, -hz);
	t[23] = -1 * nw_f(-hx, hy, hz);   t[24] = -1 * nw_f(hx, -hy, hz);
	t[25] = -1 * nw_f(hx, hy, -hz);   t[26] = -1 * nw_f(hx, hy, hz);
	return kn_sum(t, 27) / (4 * kPI * hx * hy * hz);
}

// Stored as -N (the reference's minus=true convention, Boris/DemagTFunc_Self.cpp:46-56).
inline void self_demag(double hx, double hy, double hz, double out[3])
{
	out[0] = -self_demag_1(hx, hy, hz);
	out[1] = -self_demag_1(hy, hx, hz);
	out[2] = -self_demag_1(hz, hy, hx);
}

// ---------------------------------------------------------------------------------------
// Far-field expansion of the diagonal and off-diagonal tensor elements (the OOMMF
// asymptotic series used by Boris/DemagTFunc_Asympt.cpp:8-319; credited there to
// M.J. Donahue).  The coefficient tables must be identical to the reference's, and the
// Horner grouping is kept so results agree bit-for-bit.  Each coefficient is
// scale * sum_t c_t * (hx^2)^ax (hy^2)^ay (hz^2)^az, evaluated left to right as
// ((c_t * hxpart) * hypart) * hzpart like the reference's source expressions.
struct HTerm { int c, ax, ay, az; };

struct HPows {
	double x[4], y[4], z[4];
	HPows(double hx, double hy, double hz)
	{
		x[0] = y[0] = z[0] = 1.0;
		x[1] = hx * hx; y[1] = hy * hy; z[1] = hz * hz;
		x[2] = x[1] * x[1]; y[2] = y[1] * y[1]; z[2] = z[1] * z[1];
		x[3] = x[2] * x[1]; y[3] = y[2] * y[1]; z[3] = z[2] * z[1];
	}
	double poly(const HTerm* t, int n) const
	{
		double s = 0.0;
		for (int q = 0; q < n; q++) {
			double v = (double)t[q].c;
			if (t[q].ax) v *= x[t[q].ax];
			if (t[q].ay) v *= y[t[q].ay];
			if (t[q].az) v *= z[t[q].az];
			s = (q == 0) ? v : s + v;
		}
		return s;
	}
};

// Diagonal series tables (1/R^5: a1..a6, 1/R^7: b1..b10, 1/R^9: c1..c15).
static const HTerm kDA[6][3] = {
	{{8,1,0,0},{-4,0,1,0},{-4,0,0,1}}, {{-24,1,0,0},{27,0,1,0},{-3,0,0,1}},
	{{-24,1,0,0},{-3,0,1,0},{27,0,0,1}}, {{3,1,0,0},{-4,0,1,0},{1,0,0,1}},
	{{6,1,0,0},{-3,0,1,0},{-3,0,0,1}}, {{3,1,0,0},{1,0,1,0},{-4,0,0,1}} };
#define HQ(a,b,c,d,e,f) {{a,2,0,0},{b,1,1,0},{c,1,0,1},{d,0,2,0},{e,0,1,1},{f,0,0,2}}
static const HTerm kDB[10][6] = {
	HQ(32,-40,-40,12,10,12), HQ(-240,580,20,-202,-75,22), HQ(-240,20,580,22,-75,-202),
	HQ(180,-505,55,232,-75,8), HQ(360,-450,-450,-180,900,-180), HQ(180,55,-505,8,-75,232),
	HQ(-10,30,-5,-16,10,-2), HQ(-30,55,20,8,-75,22), HQ(-30,20,55,22,-75,8), HQ(-10,-5,30,-2,10,-16) };
static const int kDBcubic[10] = { -14, 105, 105, -105, 0, -105, 7, 0, 0, 7 };
#define HC(a,b,c,d,e,f,g,h,i,j) {{a,3,0,0},{b,2,1,0},{c,2,0,1},{d,1,2,0},{e,1,1,1},{f,1,0,2},{g,0,3,0},{h,0,2,1},{i,0,1,2},{j,0,0,3}}
static const HTerm kDC[15][10] = {
	HC(384,-896,-896,672,560,672,-120,-112,-112,-120),
	HC(-5376,22624,2464,-19488,-7840,672,3705,2198,938,-345),
	HC(-5376,2464,22624,672,-7840,-19488,-345,938,2198,3705),
	HC(10080,-48720,1680,49770,-2625,-630,-10440,-1050,2100,-315),
	HC(20160,-47040,-47040,-6300,133350,-6300,7065,-26670,-26670,7065),
	HC(10080,1680,-48720,-630,-2625,49770,-315,2100,-1050,-10440),
	HC(-3360,17290,-1610,-19488,5495,-588,4848,-3136,938,-75),
	HC(-10080,32970,14070,-6300,-66675,12600,-10080,53340,-26670,3015),
	HC(-10080,14070,32970,12600,-66675,-6300,3015,-26670,53340,-10080),
	HC(-3360,-1610,17290,-588,5495,-19488,-75,938,-3136,4848),
	HC(105,-560,70,672,-280,42,-192,224,-112,15),
	HC(420,-1610,-350,672,2345,-588,528,-3136,2198,-345),
	HC(630,-1470,-1470,-630,5250,-630,360,-1050,-1050,360),
	HC(420,-350,-1610,-588,2345,672,-345,2198,-3136,528),
	HC(105,70,-560,42,-280,672,15,-112,224,-192) };
static const int kDCcubic[15] = { 32, -448, -448, -150, 7620, -150, 314, -3810, -3810, 314, -16, 134, 300, 134, -16 };

// Off-diagonal series tables (1/R^5: a1..a3, 1/R^7: b1..b6, 1/R^9: c1..c10).
static const HTerm kOA[3][3] = { {{4,1,0,0},{-3,0,1,0},{-1,0,0,1}}, {{-3,1,0,0},{4,0,1,0},{-1,0,0,1}},
	{{-3,1,0,0},{-3,0,1,0},{6,0,0,1}} };
static const HTerm kOB[6][6] = {
	HQ(16,-30,-10,10,5,2), HQ(-40,105,-5,-40,-5,4), HQ(-40,-15,115,20,-35,-32),
	HQ(10,-30,5,16,-10,2), HQ(20,-15,-35,-40,115,-32), HQ(10,15,-40,10,-40,32) };
static const int kOBcubic[6] = { -7, 19, 13, -7, 13, -13 };
static const HTerm kOC[10][10] = {
	HC(576,-2016,-672,1680,840,336,-315,-210,-126,-45),
	HC(-3024,13664,448,-12670,-2485,546,2520,910,84,-135),
	HC(-3024,1344,12768,2730,-10185,-8694,-945,1680,2394,1350),
	HC(2520,-12670,910,13664,-2485,84,-3024,448,546,-135),
	HC(5040,-9940,-13580,-9940,49700,-6300,5040,-13580,-6300,2700),
	HC(2520,2730,-14490,420,-7875,17640,-945,3990,-840,-3600),
	HC(-315,1680,-210,-2016,840,-126,576,-672,336,-45),
	HC(-945,2730,1680,1344,-10185,2394,-3024,12768,-8694,1350),
	HC(-945,420,3990,2730,-7875,-840,2520,-14490,17640,-3600),
	HC(-315,-630,2100,-630,3150,-3360,-315,2100,-3360,1440) };
static const int kOCcubic[10] = { 48, -142, -582, -142, 2840, -450, 48, -582, -450, 180 };
#undef HQ
#undef HC

struct AsymptDiag {
	bool cubic;
	double lead, a[6], b[10], c[15];
	AsymptDiag(double hx, double hy, double hz)
	{
		HPows p(hx, hy, hz);
		lead = (-hx * hy * hz / (4 * kPI));
		cubic = !(p.x[1] != p.y[1] || p.x[1] != p.z[1] || p.y[1] != p.z[1]);
		for (int q = 0; q < 6; q++) a[q] = cubic ? 0.0 : (lead / 4.0) * p.poly(kDA[q], 3);
		for (int q = 0; q < 10; q++) b[q] = cubic ? (kDBcubic[q] ? (lead / 16.0) * (kDBcubic[q] * p.x[2]) : 0.0) : (lead / 16.0) * p.poly(kDB[q], 6);
		for (int q = 0; q < 15; q++) c[q] = cubic ? (lead / 192.0) * (kDCcubic[q] * p.x[3]) : (lead / 192.0) * p.poly(kDC[q], 10);
	}
	double eval(double x, double y, double z) const
	{
		double R2 = x * x + y * y + z * z, R = std::sqrt(R2);
		if (R2 <= 0.0) return 0.0;
		double tx = x * x / (R2 * R2), ty = y * y / (R2 * R2), tz = z * z / (R2 * R2);
		double iR = 1 / R;
		double tz4 = tz * tz, tz6 = tz4 * tz;
		double t3 = (2 * tx - ty - tz) * lead, t5 = 0.0, t7;
		if (cubic) {
			double ty4 = ty * ty;
			t7 = ((b[0] * tx + (b[1] * ty + b[2] * tz)) * tx + (b[3] * ty4 + b[5] * tz4)) * tx + b[6] * ty4 * ty + b[9] * tz6;
		}
		else {
			t5 = (a[0] * tx + (a[1] * ty + a[2] * tz)) * tx + (a[3] * ty + a[4] * tz) * ty + a[5] * tz4;
			t7 = ((b[0] * tx + (b[1] * ty + b[2] * tz)) * tx + ((b[3] * ty + b[4] * tz) * ty + b[5] * tz4)) * tx
				+ ((b[6] * ty + b[7] * tz) * ty + b[8] * tz4) * ty + b[9] * tz6;
		}
		double t9 = (((c[0] * tx + (c[1] * ty + c[2] * tz)) * tx + ((c[3] * ty + c[4] * tz) * ty + c[5] * tz4)) * tx
			+ (((c[6] * ty + c[7] * tz) * ty + c[8] * tz4) * ty + c[9] * tz6)) * tx
			+ (((c[10] * ty + c[11] * tz) * ty + c[12] * tz4) * ty + c[13] * tz6) * ty + c[14] * tz4 * tz4;
		return (t9 + t7 + t5 + t3) * iR;
	}
};

struct AsymptOffDiag {
	bool cubic;
	double lead, a[3], b[6], c[10];
	AsymptOffDiag(double hx, double hy, double hz)
	{
		HPows p(hx, hy, hz);
		lead = (-hx * hy * hz / (4 * kPI));
		cubic = !(p.x[1] != p.y[1] || p.x[1] != p.z[1] || p.y[1] != p.z[1]);
		for (int q = 0; q < 3; q++) a[q] = cubic ? 0.0 : ((lead * 5.0) / 4.0) * p.poly(kOA[q], 3);
		for (int q = 0; q < 6; q++) b[q] = cubic ? ((lead * 7.0) / 16.0) * (kOBcubic[q] * p.x[2]) : ((lead * 7.0) / 16.0) * p.poly(kOB[q], 6);
		for (int q = 0; q < 10; q++) c[q] = cubic ? (lead / 64.0) * (kOCcubic[q] * p.x[3]) : (lead / 64.0) * p.poly(kOC[q], 10);
	}
	double eval(double x, double y, double z) const
	{
		double R2 = x * x + y * y + z * z, R = std::sqrt(R2);
		if (R2 <= 0.0) return 0.0;
		double tx = x * x / (R2 * R2), ty = y * y / (R2 * R2), tz = z * z / (R2 * R2);
		double iR = 1 / R, iR2 = 1 / R2;
		double t3 = 3 * lead, t5 = 0.0;
		if (!cubic) t5 = a[0] * tx + a[1] * ty + a[2] * tz;
		double tz4 = tz * tz;
		double t7 = (b[0] * tx + (b[1] * ty + b[2] * tz)) * tx + (b[3] * ty + b[4] * tz) * ty + b[5] * tz4;
		double t9 = ((c[0] * tx + (c[1] * ty + c[2] * tz)) * tx + ((c[3] * ty + c[4] * tz) * ty + c[5] * tz4)) * tx
			+ ((c[6] * ty + c[7] * tz) * ty + c[8] * tz4) * ty + c[9] * tz4 * tz;
		double iR5 = iR2 * iR2 * iR;
		return (t9 + t7 + t5 + t3) * iR5 * x * y;
	}
};

inline bool use_asympt(int i, int j, int k, int d)
{
	return d > 0 && (i >= d || j >= d || k >= d || i * i + j * j + k * k >= d * d);
}

// Index of a signed offset in a periodic axis of length N.
inline size_t wrap(int i, int N) { return (size_t)((i + N) % N); }

// Demag tensor of the zero-shift, non-periodic kernel, stored as -N, placed on the full
// N grid with the mirror symmetries (Boris/DemagTFunc.cpp:21-170 for 3D).  Ddiag and
// Dodiag are N.x*N.y*N.z DBL3, x fastest.  Off-diagonal sign patterns follow the
// reference's write order: where an index is zero the later (negated) write wins, which
// the sign table below reproduces.
inline void tensor3d(int nx, int ny, int nzc, int Nx, int Ny, int Nz, double hx, double hy, double hz,
	std::vector<double>& Dd, std::vector<double>& Do)
{
	const int d = kAsymptoticDistance;
	int dx = std::min(nx, d), dy = std::min(ny, d), dz = std::min(nzc, d);
	size_t NN = (size_t)Nx * Ny * Nz;
	Dd.assign(NN * 3, 0.0); Do.assign(NN * 3, 0.0);

	Grid3 fxx = fill_grid(nw_f, dx, dy, dz, hx, hy, hz);
	Grid3 fyy = fill_grid(nw_f, dy, dx, dz, hy, hx, hz);
	Grid3 fzz = fill_grid(nw_f, dz, dy, dx, hz, hy, hx);
	Grid3 gxy = fill_grid(nw_g, dx, dy, dz, hx, hy, hz);
	Grid3 gxz = fill_grid(nw_g, dx, dz, dy, hx, hz, hy);
	Grid3 gyz = fill_grid(nw_g, dy, dz, dx, hy, hz, hx);
	AsymptDiag adx(hx, hy, hz), ady(hy, hx, hz), adz(hz, hy, hx);
	AsymptOffDiag aoxy(hx, hy, hz), aoxz(hx, hz, hy), aoyz(hy, hz, hx);

	// (sign of i, j, k) for the 8 mirror writes, in the reference's order.
	static const int mir[8][3] = { {1,1,1},{-1,1,1},{1,-1,1},{1,1,-1},{-1,-1,1},{-1,1,-1},{1,-1,-1},{-1,-1,-1} };

#pragma omp parallel for collapse(2) schedule(dynamic, 1)
	for (int k = 0; k < nzc; k++) {
		for (int j = 0; j < ny; j++) {
			for (int i = 0; i < nx; i++) {
				double vd[3], vo[3];
				if (use_asympt(i, j, k, d)) {
					vd[0] = -adx.eval(i * hx, j * hy, k * hz);
					vd[1] = -ady.eval(j * hy, i * hx, k * hz);
					vd[2] = -adz.eval(k * hz, j * hy, i * hx);
					vo[0] = -aoxy.eval(i * hx, j * hy, k * hz);
					vo[1] = -aoxz.eval(i * hx, k * hz, j * hy);
					vo[2] = -aoyz.eval(j * hy, k * hz, i * hx);
				}
				else {
					vd[0] = -L27(fxx, i, j, k, hx, hy, hz);
					vd[1] = -L27(fyy, j, i, k, hy, hx, hz);
					vd[2] = -L27(fzz, k, j, i, hz, hy, hx);
					vo[0] = -L27(gxy, i, j, k, hx, hy, hz);
					vo[1] = -L27(gxz, i, k, j, hx, hz, hy);
					vo[2] = -L27(gyz, j, k, i, hy, hz, hx);
				}
				for (int m = 0; m < 8; m++) {
					size_t idx = wrap(mir[m][0] * i, Nx) + wrap(mir[m][1] * j, Ny) * Nx + wrap(mir[m][2] * k, Nz) * (size_t)Nx * Ny;
					// The last of the 8 writes to an index wins: a zero offset maps the
					// mirrored writes onto the same index, so emulate the overwrite order.
					int last = m;
					for (int q = m + 1; q < 8; q++) {
						if ((mir[q][0] * i == mir[m][0] * i) && (mir[q][1] * j == mir[m][1] * j) && (mir[q][2] * k == mir[m][2] * k)) last = q;
					}
					if (last != m) continue;
					int si = mir[m][0], sj = mir[m][1], sk = mir[m][2];
					Dd[idx * 3 + 0] = vd[0]; Dd[idx * 3 + 1] = vd[1]; Dd[idx * 3 + 2] = vd[2];
					Do[idx * 3 + 0] = vo[0] * (si * sj);
					Do[idx * 3 + 1] = vo[1] * (si * sk);
					Do[idx * 3 + 2] = vo[2] * (sj * sk);
				}
			}
		}
	}
}

// 2D version (n.z == 1): Boris/DemagTFunc.cpp:172-289.  Dodiag holds only Dxy.
inline void tensor2d(int nx, int ny, int Nx, int Ny, double hx, double hy, double hz,
	std::vector<double>& Dd, std::vector<double>& Dxy)
{
	const int d = kAsymptoticDistance;
	int dx = std::min(nx, d), dy = std::min(ny, d), dz = 1;
	size_t NN = (size_t)Nx * Ny;
	Dd.assign(NN * 3, 0.0); Dxy.assign(NN, 0.0);

	Grid3 fxx = fill_grid(nw_f, dx, dy, dz, hx, hy, hz);
	Grid3 fyy = fill_grid(nw_f, dy, dx, dz, hy, hx, hz);
	Grid3 fzz = fill_grid(nw_f, dz, dy, dx, hz, hy, hx);
	Grid3 gxy = fill_grid(nw_g, dx, dy, dz, hx, hy, hz);
	AsymptDiag adx(hx, hy, hz), ady(hy, hx, hz), adz(hz, hy, hx);
	AsymptOffDiag aoxy(hx, hy, hz);
	static const int mir[4][2] = { {1,1},{-1,1},{1,-1},{-1,-1} };

#pragma omp parallel for schedule(dynamic, 1)
	for (int j = 0; j < ny; j++) {
		for (int i = 0; i < nx; i++) {
			double vd[3], vo;
			if (d > 0 && (i >= d || j >= d || i * i + j * j >= d * d)) {
				vd[0] = -adx.eval(i * hx, j * hy, 0);
				vd[1] = -ady.eval(j * hy, i * hx, 0);
				vd[2] = -adz.eval(0, j * hy, i * hx);
				vo = -aoxy.eval(i * hx, j * hy, 0);
			}
			else {
				vd[0] = -L27(fxx, i, j, 0, hx, hy, hz);
				vd[1] = -L27(fyy, j, i, 0, hy, hx, hz);
				vd[2] = -L27(fzz, 0, j, i, hz, hy, hx);
				vo = -L27(gxy, i, j, 0, hx, hy, hz);
			}
			for (int m = 0; m < 4; m++) {
				int last = m;
				for (int q = m + 1; q < 4; q++)
					if ((mir[q][0] * i == mir[m][0] * i) && (mir[q][1] * j == mir[m][1] * j)) last = q;
				if (last != m) continue;
				size_t idx = wrap(mir[m][0] * i, Nx) + wrap(mir[m][1] * j, Ny) * Nx;
				Dd[idx * 3 + 0] = vd[0]; Dd[idx * 3 + 1] = vd[1]; Dd[idx * 3 + 2] = vd[2];
				Dxy[idx] = vo * (mir[m][0] * mir[m][1]);
			}
		}
	}
}

// ---------------------------------------------------------------------------------------
// Self coefficient with periodic images (DemagTFunc::SelfDemag_PBC,
// Boris/DemagTFunc_Self.cpp:58-99), stored as -N: the zero image contributes the cell's own
// self term, every other image (offset |img| n cells) the far-field series, summed in the
// reference's loop order (x images outermost).  The reference sets up all three series
// objects with (hx, hy, hz) (its yy and zz objects are not permuted); kept.
inline void self_demag_pbc(double hx, double hy, double hz, int nx, int ny, int nz, int px, int py, int pz, double out[3])
{
	AsymptDiag a(hx, hy, hz);
	double v[3] = { 0, 0, 0 };
	for (int ii = -px; ii <= px; ii++) for (int jj = -py; jj <= py; jj++) for (int kk = -pz; kk <= pz; kk++) {
		const int i = (int)std::fabs(ii * (double)nx), j = (int)std::fabs(jj * (double)ny), k = (int)std::fabs(kk * (double)nz);
		if (i == 0 && j == 0 && k == 0) {
			double s[3];
			self_demag(hx, hy, hz, s);
			for (int c = 0; c < 3; c++) v[c] += s[c];
		}
		else {
			v[0] += -a.eval(i * hx, j * hy, k * hz);
			v[1] += -a.eval(j * hy, i * hx, k * hz);
			v[2] += -a.eval(k * hz, j * hy, i * hx);
		}
	}
	for (int c = 0; c < 3; c++) out[c] = v[c];
}

// Periodic boundary conditions (Boris/DemagTFunc_PBC.cpp).  Along a periodic axis N equals
// the cell count and every offset 0..N-1 sums the tensor over the images -p..p of the
// source (offset + img*N); along an open axis N = 2n, offsets 0..n-1 are computed and their
// mirrors written.  Off-diagonal terms take sign(offset) per axis with sign(0) = 0 (the
// reference's get_sign, BorisLib/Funcs_Math.h:103), so a zero offset contributes nothing.
inline int isign0(int v) { return (v > 0) - (v < 0); }

// 3D: Boris/DemagTFunc_PBC.cpp:186-298 (diagonal) and :300-422 (off-diagonal).
inline void tensor3d_pbc(int Nx, int Ny, int Nz, double hx, double hy, double hz, int px, int py, int pz,
	std::vector<double>& Dd, std::vector<double>& Do)
{
	const int d = kAsymptoticDistance;
	px = std::abs(px); py = std::abs(py); pz = std::abs(pz);
	// f/g grids: asymptotic distance along periodic axes, n (capped at d) along open ones
	const int dx = std::min(px ? d : Nx / 2, d), dy = std::min(py ? d : Ny / 2, d), dz = std::min(pz ? d : Nz / 2, d);
	const size_t NN = (size_t)Nx * Ny * Nz;
	Dd.assign(NN * 3, 0.0); Do.assign(NN * 3, 0.0);
	Grid3 fxx = fill_grid(nw_f, dx, dy, dz, hx, hy, hz);
	Grid3 fyy = fill_grid(nw_f, dy, dx, dz, hy, hx, hz);
	Grid3 fzz = fill_grid(nw_f, dz, dy, dx, hz, hy, hx);
	Grid3 gxy = fill_grid(nw_g, dx, dy, dz, hx, hy, hz);
	Grid3 gxz = fill_grid(nw_g, dx, dz, dy, hx, hz, hy);
	Grid3 gyz = fill_grid(nw_g, dy, dz, dx, hy, hz, hx);
	AsymptDiag adx(hx, hy, hz), ady(hy, hx, hz), adz(hz, hy, hx);
	AsymptOffDiag aoxy(hx, hy, hz), aoxz(hx, hz, hy), aoyz(hy, hz, hx);
	const int ex = px ? Nx : Nx / 2, ey = py ? Ny : Ny / 2, ez = pz ? Nz : Nz / 2;

#pragma omp parallel for collapse(2) schedule(dynamic, 1)
	for (int j0 = 0; j0 < ey; j0++) {
		for (int k0 = 0; k0 < ez; k0++) {
			for (int i0 = 0; i0 < ex; i0++) {
				double vd[3] = { 0, 0, 0 }, vo[3] = { 0, 0, 0 };
				for (int ii = -px; ii <= px; ii++) for (int jj = -py; jj <= py; jj++) for (int kk = -pz; kk <= pz; kk++) {
					const int is = ii * Nx + i0, js = jj * Ny + j0, ks = kk * Nz + k0;
					const int si = isign0(is), sj = isign0(js), sk = isign0(ks);
					const int i = std::abs(is), j = std::abs(js), k = std::abs(ks);
					if (use_asympt(i, j, k, d)) {
						vd[0] += -adx.eval(i * hx, j * hy, k * hz);
						vd[1] += -ady.eval(j * hy, i * hx, k * hz);
						vd[2] += -adz.eval(k * hz, j * hy, i * hx);
						vo[0] += -aoxy.eval(i * hx, j * hy, k * hz) * (double)(si * sj);
						vo[1] += -aoxz.eval(i * hx, k * hz, j * hy) * (double)(si * sk);
						vo[2] += -aoyz.eval(j * hy, k * hz, i * hx) * (double)(sj * sk);
					}
					else {
						vd[0] += -L27(fxx, i, j, k, hx, hy, hz);
						vd[1] += -L27(fyy, j, i, k, hy, hx, hz);
						vd[2] += -L27(fzz, k, j, i, hz, hy, hx);
						vo[0] += -L27(gxy, i, j, k, hx, hy, hz) * (double)(si * sj);
						vo[1] += -L27(gxz, i, k, j, hx, hz, hy) * (double)(si * sk);
						vo[2] += -L27(gyz, j, k, i, hy, hz, hx) * (double)(sj * sk);
					}
				}
				// mirrors along open axes with a nonzero offset (no two of them coincide)
				const bool fx = !px && i0, fy = !py && j0, fz = !pz && k0;
				for (int m = 0; m < 8; m++) {
					const int mx = m & 1, my = (m >> 1) & 1, mz = (m >> 2) & 1;
					if ((mx && !fx) || (my && !fy) || (mz && !fz)) continue;
					const int sx = mx ? -1 : 1, sy = my ? -1 : 1, sz = mz ? -1 : 1;
					const size_t idx = wrap(sx * i0, Nx) + wrap(sy * j0, Ny) * Nx + wrap(sz * k0, Nz) * (size_t)Nx * Ny;
					Dd[idx * 3 + 0] = vd[0]; Dd[idx * 3 + 1] = vd[1]; Dd[idx * 3 + 2] = vd[2];
					Do[idx * 3 + 0] = vo[0] * (sx * sy); Do[idx * 3 + 1] = vo[1] * (sx * sz); Do[idx * 3 + 2] = vo[2] * (sy * sz);
				}
			}
		}
	}
}

// 2D (n.z = 1): Boris/DemagTFunc_PBC.cpp:5-94 (diagonal) and :96-184 (Dxy).  z images sit
// at whole multiples of the cell thickness.
inline void tensor2d_pbc(int Nx, int Ny, double hx, double hy, double hz, int px, int py, int pz,
	std::vector<double>& Dd, std::vector<double>& Dxy)
{
	const int d = kAsymptoticDistance;
	px = std::abs(px); py = std::abs(py); pz = std::abs(pz);
	const int dx = std::min(px ? d : Nx / 2, d), dy = std::min(py ? d : Ny / 2, d), dz = std::min(pz ? d : 1, d);
	const size_t NN = (size_t)Nx * Ny;
	Dd.assign(NN * 3, 0.0); Dxy.assign(NN, 0.0);
	Grid3 fxx = fill_grid(nw_f, dx, dy, dz, hx, hy, hz);
	Grid3 fyy = fill_grid(nw_f, dy, dx, dz, hy, hx, hz);
	Grid3 fzz = fill_grid(nw_f, dz, dy, dx, hz, hy, hx);
	Grid3 gxy = fill_grid(nw_g, dx, dy, dz, hx, hy, hz);
	AsymptDiag adx(hx, hy, hz), ady(hy, hx, hz), adz(hz, hy, hx);
	AsymptOffDiag aoxy(hx, hy, hz);
	const int ex = px ? Nx : Nx / 2, ey = py ? Ny : Ny / 2;

#pragma omp parallel for schedule(dynamic, 1)
	for (int j0 = 0; j0 < ey; j0++) {
		for (int i0 = 0; i0 < ex; i0++) {
			double vd[3] = { 0, 0, 0 }, vo = 0;
			for (int ii = -px; ii <= px; ii++) for (int jj = -py; jj <= py; jj++) for (int kk = -pz; kk <= pz; kk++) {
				const int is = ii * Nx + i0, js = jj * Ny + j0;
				const int si = isign0(is), sj = isign0(js);
				const int i = std::abs(is), j = std::abs(js), k = std::abs(kk);
				if (use_asympt(i, j, k, d)) {
					vd[0] += -adx.eval(i * hx, j * hy, k * hz);
					vd[1] += -ady.eval(j * hy, i * hx, k * hz);
					vd[2] += -adz.eval(k * hz, j * hy, i * hx);
					vo += -aoxy.eval(i * hx, j * hy, k * hz) * (double)(si * sj);
				}
				else {
					vd[0] += -L27(fxx, i, j, k, hx, hy, hz);
					vd[1] += -L27(fyy, j, i, k, hy, hx, hz);
					vd[2] += -L27(fzz, k, j, i, hz, hy, hx);
					vo += -L27(gxy, i, j, k, hx, hy, hz) * (double)(si * sj);
				}
			}
			const bool fx = !px && i0, fy = !py && j0;
			for (int m = 0; m < 4; m++) {
				const int mx = m & 1, my = (m >> 1) & 1;
				if ((mx && !fx) || (my && !fy)) continue;
				const int sx = mx ? -1 : 1, sy = my ? -1 : 1;
				const size_t idx = wrap(sx * i0, Nx) + wrap(sy * j0, Ny) * Nx;
				Dd[idx * 3 + 0] = vd[0]; Dd[idx * 3 + 1] = vd[1]; Dd[idx * 3 + 2] = vd[2];
				Dxy[idx] = vo * (sx * sy);
			}
		}
	}
}

} // namespace oracle
