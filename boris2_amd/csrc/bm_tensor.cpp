// Host-side coefficients of the far-field (asymptotic) demag-tensor series.
//
// The series and its coefficient table are the OOMMF expansion that the reference uses
// (Boris/DemagTFunc_Asympt.cpp:8-319, credited there to M.J. Donahue): 1/R^5, 1/R^7 and
// 1/R^9 terms whose coefficients are polynomials in the squared cell ratios.  The table
// has to be the same for the far field to agree; it is written here as integer
// coefficients over a fixed monomial basis and evaluated left to right with each term
// formed as ((c * hx-part) * hy-part) * hz-part, which reproduces the reference's
// floating-point results.  Cubic cells use the reference's dedicated (pre-summed)
// coefficients and Horner grouping.
#include "bm_newell.h"

namespace bm {

namespace {

// monomial bases: degree 1 (hx2, hy2, hz2); degree 2 (hx4, hx2hy2, hx2hz2, hy4, hy2hz2,
// hz4); degree 3 (hx6, hx4hy2, hx4hz2, hx2hy4, hx2hy2hz2, hx2hz4, hy6, hy4hz2, hy2hz4, hz6)
const int kDeg1[3][3] = { {1,0,0},{0,1,0},{0,0,1} };
const int kDeg2[6][3] = { {2,0,0},{1,1,0},{1,0,1},{0,2,0},{0,1,1},{0,0,2} };
const int kDeg3[10][3] = { {3,0,0},{2,1,0},{2,0,1},{1,2,0},{1,1,1},{1,0,2},{0,3,0},{0,2,1},{0,1,2},{0,0,3} };

const int kDiag5[6][3] = { {8,-4,-4},{-24,27,-3},{-24,-3,27},{3,-4,1},{6,-3,-3},{3,1,-4} };
const int kDiag7[10][6] = {
	{32,-40,-40,12,10,12}, {-240,580,20,-202,-75,22}, {-240,20,580,22,-75,-202},
	{180,-505,55,232,-75,8}, {360,-450,-450,-180,900,-180}, {180,55,-505,8,-75,232},
	{-10,30,-5,-16,10,-2}, {-30,55,20,8,-75,22}, {-30,20,55,22,-75,8}, {-10,-5,30,-2,10,-16} };
const int kDiag7Cubic[10] = { -14, 105, 105, -105, 0, -105, 7, 0, 0, 7 };
const int kDiag9[15][10] = {
	{384,-896,-896,672,560,672,-120,-112,-112,-120},
	{-5376,22624,2464,-19488,-7840,672,3705,2198,938,-345},
	{-5376,2464,22624,672,-7840,-19488,-345,938,2198,3705},
	{10080,-48720,1680,49770,-2625,-630,-10440,-1050,2100,-315},
	{20160,-47040,-47040,-6300,133350,-6300,7065,-26670,-26670,7065},
	{10080,1680,-48720,-630,-2625,49770,-315,2100,-1050,-10440},
	{-3360,17290,-1610,-19488,5495,-588,4848,-3136,938,-75},
	{-10080,32970,14070,-6300,-66675,12600,-10080,53340,-26670,3015},
	{-10080,14070,32970,12600,-66675,-6300,3015,-26670,53340,-10080},
	{-3360,-1610,17290,-588,5495,-19488,-75,938,-3136,4848},
	{105,-560,70,672,-280,42,-192,224,-112,15},
	{420,-1610,-350,672,2345,-588,528,-3136,2198,-345},
	{630,-1470,-1470,-630,5250,-630,360,-1050,-1050,360},
	{420,-350,-1610,-588,2345,672,-345,2198,-3136,528},
	{105,70,-560,42,-280,672,15,-112,224,-192} };
const int kDiag9Cubic[15] = { 32, -448, -448, -150, 7620, -150, 314, -3810, -3810, 314, -16, 134, 300, 134, -16 };

const int kOff5[3][3] = { {4,-3,-1},{-3,4,-1},{-3,-3,6} };
const int kOff7[6][6] = {
	{16,-30,-10,10,5,2}, {-40,105,-5,-40,-5,4}, {-40,-15,115,20,-35,-32},
	{10,-30,5,16,-10,2}, {20,-15,-35,-40,115,-32}, {10,15,-40,10,-40,32} };
const int kOff7Cubic[6] = { -7, 19, 13, -7, 13, -13 };
const int kOff9[10][10] = {
	{576,-2016,-672,1680,840,336,-315,-210,-126,-45},
	{-3024,13664,448,-12670,-2485,546,2520,910,84,-135},
	{-3024,1344,12768,2730,-10185,-8694,-945,1680,2394,1350},
	{2520,-12670,910,13664,-2485,84,-3024,448,546,-135},
	{5040,-9940,-13580,-9940,49700,-6300,5040,-13580,-6300,2700},
	{2520,2730,-14490,420,-7875,17640,-945,3990,-840,-3600},
	{-315,1680,-210,-2016,840,-126,576,-672,336,-45},
	{-945,2730,1680,1344,-10185,2394,-3024,12768,-8694,1350},
	{-945,420,3990,2730,-7875,-840,2520,-14490,17640,-3600},
	{-315,-630,2100,-630,3150,-3360,-315,2100,-3360,1440} };
const int kOff9Cubic[10] = { 48, -142, -582, -142, 2840, -450, 48, -582, -450, 180 };

struct Pw {
	double x[4], y[4], z[4];
	Pw(double hx, double hy, double hz)
	{
		x[0] = y[0] = z[0] = 1; x[1] = hx * hx; y[1] = hy * hy; z[1] = hz * hz;
		x[2] = x[1] * x[1]; y[2] = y[1] * y[1]; z[2] = z[1] * z[1];
		x[3] = x[2] * x[1]; y[3] = y[2] * y[1]; z[3] = z[2] * z[1];
	}
	bool cubic() const { return !(x[1] != y[1] || x[1] != z[1] || y[1] != z[1]); }
	template <int NT>
	double eval(const int* coef, const int (*basis)[3]) const
	{
		double s = 0;
		for (int q = 0; q < NT; q++) {
			double v = (double)coef[q];
			if (basis[q][0]) v *= x[basis[q][0]];
			if (basis[q][1]) v *= y[basis[q][1]];
			if (basis[q][2]) v *= z[basis[q][2]];
			s = q ? s + v : v;
		}
		return s;
	}
};

} // namespace

void asym_diag_init(AsymDiag& A, double hx, double hy, double hz)
{
	Pw p(hx, hy, hz);
	A.lead = (-hx * hy * hz / (4 * kPi));
	A.cubic = p.cubic() ? 1 : 0;
	for (int q = 0; q < 6; q++) A.a[q] = A.cubic ? 0.0 : (A.lead / 4.0) * p.eval<3>(kDiag5[q], kDeg1);
	for (int q = 0; q < 10; q++)
		A.b[q] = A.cubic ? (kDiag7Cubic[q] ? (A.lead / 16.0) * (kDiag7Cubic[q] * p.x[2]) : 0.0) : (A.lead / 16.0) * p.eval<6>(kDiag7[q], kDeg2);
	for (int q = 0; q < 15; q++)
		A.c[q] = A.cubic ? (A.lead / 192.0) * (kDiag9Cubic[q] * p.x[3]) : (A.lead / 192.0) * p.eval<10>(kDiag9[q], kDeg3);
}

// Self-demag coefficient of one cell, stored as -N: the 27-term second difference of f at
// zero offset (Boris/DemagTFunc_Self.cpp:6-56).  The reference evaluates it with the cell
// size in metres for the demag evaluation speedup (Boris/DemagCUDA.cpp:37); callers pass
// the same values so the zero guards of f act identically.
static double self_demag_1(double hx, double hy, double hz)
{
	const double t[27] = {
		+8 * newell_f(0, 0, 0),
		-4 * newell_f(hx, 0, 0), -4 * newell_f(-hx, 0, 0), -4 * newell_f(0, hy, 0), -4 * newell_f(0, -hy, 0),
		-4 * newell_f(0, 0, hz), -4 * newell_f(0, 0, -hz),
		+2 * newell_f(-hx, -hy, 0), +2 * newell_f(-hx, hy, 0), +2 * newell_f(hx, -hy, 0), +2 * newell_f(hx, hy, 0),
		+2 * newell_f(-hx, 0, -hz), +2 * newell_f(-hx, 0, hz), +2 * newell_f(hx, 0, -hz), +2 * newell_f(hx, 0, hz),
		+2 * newell_f(0, -hy, -hz), +2 * newell_f(0, -hy, hz), +2 * newell_f(0, hy, -hz), +2 * newell_f(0, hy, hz),
		-1 * newell_f(-hx, -hy, -hz), -1 * newell_f(-hx, -hy, hz), -1 * newell_f(-hx, hy, -hz), -1 * newell_f(hx, -hy, -hz),
		-1 * newell_f(-hx, hy, hz), -1 * newell_f(hx, -hy, hz), -1 * newell_f(hx, hy, -hz), -1 * newell_f(hx, hy, hz) };
	return comp_sum(t, 27) / (4 * kPi * hx * hy * hz);
}

void self_demag_coeff(double hx, double hy, double hz, double out[3])
{
	out[0] = -self_demag_1(hx, hy, hz);
	out[1] = -self_demag_1(hy, hx, hz);
	out[2] = -self_demag_1(hz, hy, hx);
}

// Self coefficient with periodic images (DemagTFunc::SelfDemag_PBC,
// Boris/DemagTFunc_Self.cpp:58-99), stored as -N: the zero image gives the cell's own term,
// every other image (offset |img| n cells) the far-field series, summed with x images
// outermost.  The reference sets up its three series objects all with (hx, hy, hz); the one
// object here reproduces that.
void self_demag_coeff_pbc(double hx, double hy, double hz, int nx, int ny, int nz, int px, int py, int pz, double out[3])
{
	AsymDiag A;
	asym_diag_init(A, hx, hy, hz);
	double v[3] = { 0.0, 0.0, 0.0 };
	for (int ii = -px; ii <= px; ii++)
		for (int jj = -py; jj <= py; jj++)
			for (int kk = -pz; kk <= pz; kk++) {
				const int i = (int)std::fabs(ii * (double)nx), j = (int)std::fabs(jj * (double)ny), k = (int)std::fabs(kk * (double)nz);
				if (i == 0 && j == 0 && k == 0) {
					double s[3];
					self_demag_coeff(hx, hy, hz, s);
					for (int c = 0; c < 3; c++) v[c] += s[c];
				}
				else {
					v[0] += -asym_diag(A, i * hx, j * hy, k * hz);
					v[1] += -asym_diag(A, j * hy, i * hx, k * hz);
					v[2] += -asym_diag(A, k * hz, j * hy, i * hx);
				}
			}
	for (int c = 0; c < 3; c++) out[c] = v[c];
}

void asym_off_init(AsymOff& A, double hx, double hy, double hz)
{
	Pw p(hx, hy, hz);
	A.lead = (-hx * hy * hz / (4 * kPi));
	A.cubic = p.cubic() ? 1 : 0;
	for (int q = 0; q < 3; q++) A.a[q] = A.cubic ? 0.0 : ((A.lead * 5.0) / 4.0) * p.eval<3>(kOff5[q], kDeg1);
	for (int q = 0; q < 6; q++)
		A.b[q] = A.cubic ? ((A.lead * 7.0) / 16.0) * (kOff7Cubic[q] * p.x[2]) : ((A.lead * 7.0) / 16.0) * p.eval<6>(kOff7[q], kDeg2);
	for (int q = 0; q < 10; q++)
		A.c[q] = A.cubic ? (A.lead / 64.0) * (kOff9Cubic[q] * p.x[3]) : (A.lead / 64.0) * p.eval<10>(kOff9[q], kDeg3);
}

} // namespace bm
