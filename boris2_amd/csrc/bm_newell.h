// Newell demag-tensor functions in FP64, shared by host (coefficient setup) and device
// (tensor fill).  The tensor is stored as -N (Boris's convention) and the far field uses
// the OOMMF asymptotic series with the reference's coefficient table and switch-over
// distance of 40 cells.
//
// Reference: basis functions Boris/DemagTFunc_fg.cpp:10-92; 27-point stencil
// Boris/DemagTFunc_L.cpp:8-110; asymptotic series Boris/DemagTFunc_Asympt.cpp:8-319;
// tensor placement Boris/DemagTFunc.cpp:21-289; summation BorisLib/Funcs_Math.h:294.
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>

namespace bm {

#define BM_HD __host__ __device__ __forceinline__

constexpr double kPi = 3.1415926535897932384626433833;
constexpr int kAsymDist = 40;
// zero threshold: the reference's float literal 1e-12F widened to double
#define BM_EPSZ ((double)1e-12F)

BM_HD bool nzero(double v) { return v > BM_EPSZ || v < -BM_EPSZ; }

// compensated (Kahan-Neumaier) sum in the given order
BM_HD double comp_sum(const double* v, int n)
{
	double s = 0.0, c = 0.0;
	for (int i = 0; i < n; i++) {
		double t = s + v[i];
		c += (fabs(s) > fabs(v[i])) ? ((s - t) + v[i]) : ((v[i] - t) + s);
		s = t;
	}
	return s + c;
}

BM_HD double newell_f(double x, double y, double z)
{
	double x2 = x * x, y2 = y * y, z2 = z * z, R = sqrt(x2 + y2 + z2);
	double t[4];
	t[0] = (2.0 * x2 - y2 - z2) * R / 6.0;
	t[1] = 0.0; t[2] = 0.0; t[3] = 0.0;
	double pxz = x2 + z2, pxy = x2 + y2;
	if (nzero(pxz)) { double a = 2.0 * y * (y + R) / pxz; if (a > -1.0) t[1] = y * (z2 - x2) * log1p(a) / 4.0; }
	if (nzero(pxy)) { double a = 2.0 * z * (z + R) / pxy; if (a > -1.0) t[2] = z * (y2 - x2) * log1p(a) / 4.0; }
	if (nzero(x)) t[3] = -x * y * z * atan(y * z / (x * R));
	return comp_sum(t, 4);
}

BM_HD double newell_g(double x, double y, double z)
{
	double x2 = x * x, y2 = y * y, z2 = z * z, R = sqrt(x2 + y2 + z2);
	double t[7];
	t[0] = -x * y * R / 3.0;
	t[1] = t[2] = t[3] = t[4] = t[5] = t[6] = 0.0;
	double pxy = x2 + y2, pyz = y2 + z2, pxz = x2 + z2;
	if (nzero(pxy)) { double a = 2.0 * z * (z + R) / pxy; if (a > -1.0) t[1] = x * y * z * log1p(a) / 2.0; }
	if (nzero(pyz)) { double a = 2.0 * x * (x + R) / pyz; if (a > -1.0) t[2] = y * (3.0 * z2 - y2) * log1p(a) / 12.0; }
	if (nzero(pxz)) { double a = 2.0 * y * (y + R) / pxz; if (a > -1.0) t[3] = x * (3.0 * z2 - x2) * log1p(a) / 12.0; }
	if (nzero(z)) t[4] = -z2 * z * atan(x * y / (z * R)) / 6.0;
	if (nzero(y)) t[5] = -y2 * z * atan(x * z / (y * R)) / 2.0;
	if (nzero(x)) t[6] = -x2 * z * atan(y * z / (x * R)) / 2.0;
	return comp_sum(t, 7);
}

// Second difference over the 3x3x3 neighbourhood of grid point (i,j,k) of a basis-function
// grid g (offsets -1..d, stored with +1, first index fastest, dims gx, gy).
BM_HD double stencil27(const double* g, int gx, int gy, int i, int j, int k, double hx, double hy, double hz)
{
	i++; j++; k++;
	auto G = [&](int a, int b, int c) { return g[(size_t)a + (size_t)b * gx + (size_t)c * gx * gy]; };
	double t[27] = {
		+8 * G(i, j, k),
		-4 * G(i + 1, j, k), -4 * G(i - 1, j, k), -4 * G(i, j + 1, k), -4 * G(i, j - 1, k), -4 * G(i, j, k + 1), -4 * G(i, j, k - 1),
		+2 * G(i - 1, j - 1, k), +2 * G(i - 1, j + 1, k), +2 * G(i + 1, j - 1, k), +2 * G(i + 1, j + 1, k),
		+2 * G(i - 1, j, k - 1), +2 * G(i - 1, j, k + 1), +2 * G(i + 1, j, k - 1), +2 * G(i + 1, j, k + 1),
		+2 * G(i, j - 1, k - 1), +2 * G(i, j - 1, k + 1), +2 * G(i, j + 1, k - 1), +2 * G(i, j + 1, k + 1),
		-1 * G(i - 1, j - 1, k - 1), -1 * G(i - 1, j - 1, k + 1), -1 * G(i - 1, j + 1, k - 1), -1 * G(i + 1, j - 1, k - 1),
		-1 * G(i - 1, j + 1, k + 1), -1 * G(i + 1, j - 1, k + 1), -1 * G(i + 1, j + 1, k - 1), -1 * G(i + 1, j + 1, k + 1) };
	return comp_sum(t, 27) / (4 * kPi * hx * hy * hz);
}

// ---------------------------------------------------------------------------------------
// Far-field series.  Coefficients are filled on the host (identical table to the
// reference, evaluated in the reference's operation order) and passed to the device.
struct AsymDiag { int cubic; double lead, a[6], b[10], c[15]; };
struct AsymOff { int cubic; double lead, a[3], b[6], c[10]; };

// host-side construction (bm_tensor.cpp)
void asym_diag_init(AsymDiag& A, double hx, double hy, double hz);
void asym_off_init(AsymOff& A, double hx, double hy, double hz);
// self-demag coefficient (-N diagonal) of one cell of size hx x hy x hz (bm_tensor.cpp)
void self_demag_coeff(double hx, double hy, double hz, double out[3]);
void self_demag_coeff_pbc(double hx, double hy, double hz, int nx, int ny, int nz, int px, int py, int pz, double out[3]);

BM_HD double asym_diag(const AsymDiag& A, double x, double y, double z)
{
	double R2 = x * x + y * y + z * z;
	if (R2 <= 0.0) return 0.0;
	double R = sqrt(R2);
	double tx = x * x / (R2 * R2), ty = y * y / (R2 * R2), tz = z * z / (R2 * R2);
	double iR = 1 / R, tz4 = tz * tz, tz6 = tz4 * tz;
	double t3 = (2 * tx - ty - tz) * A.lead, t5 = 0.0, t7;
	const double* b = A.b; const double* a = A.a; const double* c = A.c;
	if (A.cubic) {
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

BM_HD double asym_off(const AsymOff& A, double x, double y, double z)
{
	double R2 = x * x + y * y + z * z;
	if (R2 <= 0.0) return 0.0;
	double R = sqrt(R2);
	double tx = x * x / (R2 * R2), ty = y * y / (R2 * R2), tz = z * z / (R2 * R2);
	double iR = 1 / R, iR2 = 1 / R2;
	const double* b = A.b; const double* a = A.a; const double* c = A.c;
	double t3 = 3 * A.lead, t5 = A.cubic ? 0.0 : (a[0] * tx + a[1] * ty + a[2] * tz);
	double tz4 = tz * tz;
	double t7 = (b[0] * tx + (b[1] * ty + b[2] * tz)) * tx + (b[3] * ty + b[4] * tz) * ty + b[5] * tz4;
	double t9 = ((c[0] * tx + (c[1] * ty + c[2] * tz)) * tx + ((c[3] * ty + c[4] * tz) * ty + c[5] * tz4)) * tx
		+ ((c[6] * ty + c[7] * tz) * ty + c[8] * tz4) * ty + c[9] * tz4 * tz;
	return (t9 + t7 + t5 + t3) * (iR2 * iR2 * iR) * x * y;
}

} // namespace bm
