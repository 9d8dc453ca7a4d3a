// ORACLE — TEST INFRASTRUCTURE ONLY.
// CPU restatement of the reference's demag convolution, effective fields, LLG and the
// RKDP54 / RK4 / Heun (TEuler) / AHeun integrators with their step control.  Templated on
// the real type: FP64 is the oracle; FP32 is the "same algorithm in single precision"
// comparison the reference lacks (it has no FP32 CPU build).  Only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg may load or call anything
// under oracle/.
//
// Parity status (pins live in tests/test_oracle_*.py):
//   Fft               pinned against numpy.fft (FP64, all lengths used) and a brute DFT.
//   Kernel / Demag    pinned by: the reference tensor (oracle/_ref golden vectors) feeding
//                     a direct O(n^2) real-space sum on small meshes (FFT path == direct
//                     sum to 1e-12), the Aharoni closed form for a uniformly magnetised
//                     prism, and the reference-held vortex OVF2 states (input only).
//                     Kernel values themselves (the FFTW boundary of the reference) are
//                     "parity unpinned" at bit level; pinned only through the field.
//   Exchange          pinned by closed forms: uniform M -> 0, cos(kx) eigenvalue.
//   Anisotropy/Zeeman pinned by closed forms (single-cell values).
//   LLG, RK stages    pinned by the macrospin closed form and order-of-convergence tests
//                     of each tableau on scalar problems.
//   Step control      closed formula (tests replay SetAdaptiveTimeStep by hand).
//   BorisRand         textbook LCG sequence (Numerical Recipes constants).
//   Eval speedup      Lagrange coefficients reproduce polynomials of degree speedup-1
//                     exactly; slot choice keeps the last `speedup` computed fields; a
//                     one-cell mesh (demag = self term only) integrates identically with
//                     and without the speedup; trajectories converge onto the plain ones
//                     as dT shrinks.
#pragma once

#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <vector>
#include <functional>
#include <algorithm>
#include <stdexcept>

#include "oracle_tensor.hpp"

namespace oracle {

// OpenMP regions run in parallel only above this many cells (or padded convolution points):
// below it the fork and join of the thread team costs more than the work (one-cell and
// macrospin runs make millions of such regions).
constexpr size_t kOmpMinWork = (size_t)1 << 14;

static const double kMU0 = 1.256637061435916e-6;   // BorisLib/Funcs_Math_base.h:9
static const double kGAMMA = 2.212761569e5;        // BorisLib/Funcs_Math_base.h:16

// ---------------------------------------------------------------------------------------
// Complex FFT of arbitrary length: Stockham autosort with radices 4, 2, 3, 5 and a direct
// DFT for any remaining prime factor.  Twiddles are computed in FP64 and rounded to R.
// sign = -1 forward (exp(-i...)), +1 inverse, unnormalised.
template <typename R>
struct Fft {
	typedef std::complex<R> C;
	int n = 0;
	std::vector<int> radix;
	std::vector<C> twf, twi;   // exp(-2 pi i m / n), exp(+2 pi i m / n)

	Fft() {}
	explicit Fft(int n_) : n(n_)
	{
		int m = n;
		while (m % 4 == 0) { radix.push_back(4); m /= 4; }
		while (m % 2 == 0) { radix.push_back(2); m /= 2; }
		while (m % 3 == 0) { radix.push_back(3); m /= 3; }
		while (m % 5 == 0) { radix.push_back(5); m /= 5; }
		for (int p = 7; m > 1; p += 2) while (m % p == 0) { radix.push_back(p); m /= p; }
		twf.resize(n); twi.resize(n);
		for (int k = 0; k < n; k++) {
			long double a = 2.0L * 3.14159265358979323846264338327950288L * (long double)k / (long double)n;
			twf[k] = C((R)std::cos((double)a), (R)-std::sin((double)a));
			twi[k] = C((R)std::cos((double)a), (R)std::sin((double)a));
		}
	}

	// In place on x (length n) using scratch y (length n).
	void run(C* x, C* y, int sign) const
	{
		const std::vector<C>& tw = (sign < 0) ? twf : twi;
		C* in = x; C* out = y;
		int Ns = 1;
		for (int r : radix) {
			int nb = n / r;
			for (int j = 0; j < nb; j++) {
				int k = j % Ns;
				C v[64]; C* vv = v; std::vector<C> big;
				if (r > 64) { big.resize(r); vv = big.data(); }
				for (int q = 0; q < r; q++) {
					C a = in[j + q * nb];
					if (k && q) a *= tw[(size_t)((long long)k * q * (n / (Ns * r)) % n)];
					vv[q] = a;
				}
				int base = (j / Ns) * Ns * r + k;
				if (r == 2) {
					C a = vv[0], b = vv[1];
					out[base] = a + b; out[base + Ns] = a - b;
				}
				else if (r == 4) {
					C a0 = vv[0] + vv[2], a1 = vv[0] - vv[2], a2 = vv[1] + vv[3], a3 = vv[1] - vv[3];
					C ja3 = (sign < 0) ? C(a3.imag(), -a3.real()) : C(-a3.imag(), a3.real());
					out[base] = a0 + a2; out[base + Ns] = a1 + ja3; out[base + 2 * Ns] = a0 - a2; out[base + 3 * Ns] = a1 - ja3;
				}
				else {
					// direct DFT of length r
					for (int p = 0; p < r; p++) {
						C s(0, 0);
						for (int q = 0; q < r; q++) s += vv[q] * tw[(size_t)(((long long)p * q % r) * (n / r))];
						out[base + p * Ns] = s;
					}
				}
			}
			std::swap(in, out);
			Ns *= r;
		}
		if (in != x) std::memcpy(x, in, sizeof(C) * n);
	}
};

// ---------------------------------------------------------------------------------------
// Demag kernels from the tensor: per-axis FFTs of the full real tensor, keeping the real
// or imaginary part the axis parity dictates (Boris/DemagKernel.cpp:171-516).  Output
// layout: Kdiag/Kodiag (3 doubles per point) on (Nx/2+1) x (Ny/2+1) x (Nz/2+1), x fastest;
// 2D: Kdiag (3) and K2D_odiag (1) on (Nx/2+1) x (Ny/2+1).
struct Kernels {
	int nx, ny, nz, Nx, Ny, Nz;
	double hx, hy, hz;
	std::vector<double> Kdiag, Kodiag;   // 3D: 3 per point each; 2D: Kdiag 3, Kodiag 1
};

inline void kernels_from_tensor_3d(Kernels& K, std::vector<double>& D, bool offdiag, std::vector<double>& out)
{
	int Nx = K.Nx, Ny = K.Ny, Nz = K.Nz, Nc = Nx / 2 + 1, Nyh = Ny / 2 + 1, Nzh = Nz / 2 + 1;
	typedef std::complex<double> C;
	// x: real lines of length Nx -> keep Nc; stored back with row stride Nc
	{
		Fft<double> f(Nx);
		std::vector<double> tmp((size_t)Nc * Ny * Nz * 3);
#pragma omp parallel if((size_t)Nx * Ny * Nz >= kOmpMinWork)
		{
			std::vector<C> a(Nx), s(Nx);
#pragma omp for collapse(2)
			for (int k = 0; k < Nz; k++) for (int j = 0; j < Ny; j++) {
				for (int c = 0; c < 3; c++) {
					for (int i = 0; i < Nx; i++) a[i] = C(D[((size_t)i + (size_t)j * Nx + (size_t)k * Nx * Ny) * 3 + c], 0.0);
					f.run(a.data(), s.data(), -1);
					for (int i = 0; i < Nc; i++) {
						double v = (!offdiag || c == 2) ? a[i].real() : a[i].imag();
						tmp[((size_t)i + (size_t)j * Nc + (size_t)k * Nc * Ny) * 3 + c] = v;
					}
				}
			}
		}
		D.swap(tmp);
	}
	// y: lines of length Ny -> keep Nyh
	{
		Fft<double> f(Ny);
		std::vector<double> tmp((size_t)Nc * Nyh * Nz * 3);
#pragma omp parallel if((size_t)Nx * Ny * Nz >= kOmpMinWork)
		{
			std::vector<C> a(Ny), s(Ny);
#pragma omp for collapse(2)
			for (int k = 0; k < Nz; k++) for (int i = 0; i < Nc; i++) {
				for (int c = 0; c < 3; c++) {
					for (int j = 0; j < Ny; j++) a[j] = C(D[((size_t)i + (size_t)j * Nc + (size_t)k * Nc * Ny) * 3 + c], 0.0);
					f.run(a.data(), s.data(), -1);
					for (int j = 0; j < Nyh; j++) {
						double v = (!offdiag || c == 1) ? a[j].real() : a[j].imag();
						tmp[((size_t)i + (size_t)j * Nc + (size_t)k * Nc * Nyh) * 3 + c] = v;
					}
				}
			}
		}
		D.swap(tmp);
	}
	// z: lines of length Nz -> keep Nzh; off-diagonal final sign -1 (i*i = -1)
	{
		Fft<double> f(Nz);
		out.assign((size_t)Nc * Nyh * Nzh * 3, 0.0);
#pragma omp parallel if((size_t)Nx * Ny * Nz >= kOmpMinWork)
		{
			std::vector<C> a(Nz), s(Nz);
#pragma omp for collapse(2)
			for (int j = 0; j < Nyh; j++) for (int i = 0; i < Nc; i++) {
				for (int c = 0; c < 3; c++) {
					for (int k = 0; k < Nz; k++) a[k] = C(D[((size_t)i + (size_t)j * Nc + (size_t)k * Nc * Nyh) * 3 + c], 0.0);
					f.run(a.data(), s.data(), -1);
					for (int k = 0; k < Nzh; k++) {
						double v;
						if (!offdiag) v = a[k].real();
						else v = (c == 0) ? -a[k].real() : -a[k].imag();
						out[((size_t)i + (size_t)j * Nc + (size_t)k * Nc * Nyh) * 3 + c] = v;
					}
				}
			}
		}
	}
}

// pbc[3]: periodic images per axis (null or all zero: open boundaries); along a periodic
// axis N equals the cell count (Boris/ConvolutionDataCUDA.cpp:56-90).
inline Kernels make_kernels(int nx, int ny, int nz, int Nx, int Ny, int Nz, double hx, double hy, double hz,
	const int* pbc = nullptr)
{
	const bool periodic = pbc && (pbc[0] || pbc[1] || pbc[2]);
	Kernels K{ nx, ny, nz, Nx, Ny, Nz, hx, hy, hz, {}, {} };
	double hm = std::max(hx, std::max(hy, hz));
	double rx = hx / hm, ry = hy / hm, rz = hz / hm;
	int Nc = Nx / 2 + 1, Nyh = Ny / 2 + 1;
	typedef std::complex<double> C;
	if (nz == 1) {
		std::vector<double> Dd, Dxy;
		if (periodic) tensor2d_pbc(Nx, Ny, rx, ry, rz, pbc[0], pbc[1], pbc[2], Dd, Dxy);
		else tensor2d(nx, ny, Nx, Ny, rx, ry, rz, Dd, Dxy);
		Fft<double> fx(Nx), fy(Ny);
		std::vector<double> t3((size_t)Nc * Ny * 3), t1((size_t)Nc * Ny);
		std::vector<C> a(std::max(Nx, Ny)), s(std::max(Nx, Ny));
		for (int j = 0; j < Ny; j++) {
			for (int c = 0; c < 4; c++) {
				for (int i = 0; i < Nx; i++) a[i] = C(c < 3 ? Dd[((size_t)i + (size_t)j * Nx) * 3 + c] : Dxy[(size_t)i + (size_t)j * Nx], 0.0);
				fx.run(a.data(), s.data(), -1);
				for (int i = 0; i < Nc; i++) {
					if (c < 3) t3[((size_t)i + (size_t)j * Nc) * 3 + c] = a[i].real();
					else t1[(size_t)i + (size_t)j * Nc] = a[i].imag();
				}
			}
		}
		K.Kdiag.assign((size_t)Nc * Nyh * 3, 0.0); K.Kodiag.assign((size_t)Nc * Nyh, 0.0);
		for (int i = 0; i < Nc; i++) {
			for (int c = 0; c < 4; c++) {
				for (int j = 0; j < Ny; j++) a[j] = C(c < 3 ? t3[((size_t)i + (size_t)j * Nc) * 3 + c] : t1[(size_t)i + (size_t)j * Nc], 0.0);
				fy.run(a.data(), s.data(), -1);
				for (int j = 0; j < Nyh; j++) {
					if (c < 3) K.Kdiag[((size_t)i + (size_t)j * Nc) * 3 + c] = a[j].real();
					else K.Kodiag[(size_t)i + (size_t)j * Nc] = -a[j].imag();
				}
			}
		}
	}
	else {
		std::vector<double> Dd, Do;
		if (periodic) tensor3d_pbc(Nx, Ny, Nz, rx, ry, rz, pbc[0], pbc[1], pbc[2], Dd, Do);
		else tensor3d(nx, ny, nz, Nx, Ny, Nz, rx, ry, rz, Dd, Do);
		kernels_from_tensor_3d(K, Dd, false, K.Kdiag);
		std::vector<double>().swap(Dd);
		kernels_from_tensor_3d(K, Do, true, K.Kodiag);
	}
	return K;
}

// ---------------------------------------------------------------------------------------
// FFT demag convolution: H = IFFT(K . FFT(M)) / (Nx Ny Nz), zero-padded planes skipped as
// in Boris/Convolution.h:262-841; kernel multiplication with the reference's symmetry
// rules (Boris/DemagKernel.cpp:42-168).
template <typename R>
struct Demag {
	typedef std::complex<R> C;
	int nx, ny, nz, Nx, Ny, Nz, Nc, Nyh, Nzh;
	std::vector<R> Kd, Ko;         // kernels rounded to R
	Fft<R> fx, fy, fz;
	std::vector<C> F[3];           // Nc x Ny x Nz spectra

	// pbc: periodic images per axis (null: open boundaries); a periodic axis has N = n
	Demag(int nx_, int ny_, int nz_, double hx, double hy, double hz, int Nz_override = 0, const int* pbc = nullptr)
		: nx(nx_), ny(ny_), nz(nz_)
	{
		const int p0 = pbc ? pbc[0] : 0, p1 = pbc ? pbc[1] : 0, p2 = pbc ? pbc[2] : 0;
		Nx = p0 ? nx : 2 * nx; Ny = p1 ? ny : 2 * ny;
		Nz = (nz == 1) ? 1 : p2 ? nz : (Nz_override ? Nz_override : 2 * nz);
		Nc = Nx / 2 + 1; Nyh = Ny / 2 + 1; Nzh = Nz / 2 + 1;
		Kernels K = make_kernels(nx, ny, nz, Nx, Ny, Nz, hx, hy, hz, pbc);
		Kd.assign(K.Kdiag.begin(), K.Kdiag.end());
		Ko.assign(K.Kodiag.begin(), K.Kodiag.end());
		fx = Fft<R>(Nx); fy = Fft<R>(Ny); if (Nz > 1) fz = Fft<R>(Nz);
		for (int c = 0; c < 3; c++) F[c].assign((size_t)Nc * Ny * Nz, C(0, 0));
	}

	size_t fidx(int i, int j, int k) const { return (size_t)i + (size_t)j * Nc + (size_t)k * Nc * Ny; }

	// M, H: n cells of 3 reals, x fastest.  clear_out: H = Hd else H += Hd.
	void convolute(const R* M, R* H, bool clear_out)
	{
		// forward x (R2C) for the nonzero rows
#pragma omp parallel if((size_t)Nx * Ny * Nz >= kOmpMinWork)
		{
			std::vector<C> a(Nx), s(Nx);
#pragma omp for collapse(2)
			for (int k = 0; k < nz; k++) for (int j = 0; j < ny; j++) {
				for (int c = 0; c < 3; c++) {
					for (int i = 0; i < Nx; i++) a[i] = C(i < nx ? M[((size_t)i + (size_t)j * nx + (size_t)k * nx * ny) * 3 + c] : (R)0, (R)0);
					fx.run(a.data(), s.data(), -1);
					for (int i = 0; i < Nc; i++) F[c][fidx(i, j, k)] = a[i];
				}
			}
		}
		// forward y for the nonzero planes
#pragma omp parallel if((size_t)Nx * Ny * Nz >= kOmpMinWork)
		{
			std::vector<C> a(Ny), s(Ny);
#pragma omp for collapse(2)
			for (int k = 0; k < nz; k++) for (int i = 0; i < Nc; i++) {
				for (int c = 0; c < 3; c++) {
					for (int j = 0; j < Ny; j++) a[j] = (j < ny) ? F[c][fidx(i, j, k)] : C(0, 0);
					fy.run(a.data(), s.data(), -1);
					for (int j = 0; j < Ny; j++) F[c][fidx(i, j, k)] = a[j];
				}
			}
		}
		if (Nz == 1) {
			// 2D: multiply along y lines
#pragma omp parallel for if((size_t)Nx * Ny * Nz >= kOmpMinWork)
			for (int i = 0; i < Nc; i++) {
				for (int j = 0; j < Ny; j++) {
					int jj = (j <= Ny / 2) ? j : Ny - j;
					R sg = (j <= Ny / 2) ? (R)1 : (R)-1;
					size_t ki = (size_t)i + (size_t)jj * Nc;
					R kxx = Kd[ki * 3 + 0], kyy = Kd[ki * 3 + 1], kzz = Kd[ki * 3 + 2], kxy = sg * Ko[ki];
					size_t f = fidx(i, j, 0);
					C mx = F[0][f], my = F[1][f], mz = F[2][f];
					F[0][f] = kxx * mx + kxy * my;
					F[1][f] = kxy * mx + kyy * my;
					F[2][f] = kzz * mz;
				}
			}
		}
		else {
#pragma omp parallel if((size_t)Nx * Ny * Nz >= kOmpMinWork)
			{
				std::vector<C> a[3], s(Nz);
				for (int c = 0; c < 3; c++) a[c].resize(Nz);
#pragma omp for collapse(2)
				for (int j = 0; j < Ny; j++) for (int i = 0; i < Nc; i++) {
					for (int c = 0; c < 3; c++) {
						for (int k = 0; k < Nz; k++) a[c][k] = (k < nz) ? F[c][fidx(i, j, k)] : C(0, 0);
						fz.run(a[c].data(), s.data(), -1);
					}
					int jj = (j <= Ny / 2) ? j : Ny - j;
					R sy = (j <= Ny / 2) ? (R)1 : (R)-1;
					for (int k = 0; k < Nz; k++) {
						int kk = (k <= Nz / 2) ? k : Nz - k;
						R sz = (k <= Nz / 2) ? (R)1 : (R)-1;
						size_t ki = (size_t)i + (size_t)jj * Nc + (size_t)kk * Nc * Nyh;
						R kxx = Kd[ki * 3 + 0], kyy = Kd[ki * 3 + 1], kzz = Kd[ki * 3 + 2];
						R kxy = sy * Ko[ki * 3 + 0], kxz = sz * Ko[ki * 3 + 1], kyz = sy * sz * Ko[ki * 3 + 2];
						C mx = a[0][k], my = a[1][k], mz = a[2][k];
						a[0][k] = kxx * mx + kxy * my + kxz * mz;
						a[1][k] = kxy * mx + kyy * my + kyz * mz;
						a[2][k] = kxz * mx + kyz * my + kzz * mz;
					}
					for (int c = 0; c < 3; c++) {
						fz.run(a[c].data(), s.data(), +1);
						for (int k = 0; k < nz; k++) F[c][fidx(i, j, k)] = a[c][k];
					}
				}
			}
		}
		// inverse y for the nonzero planes, keep j < ny
#pragma omp parallel if((size_t)Nx * Ny * Nz >= kOmpMinWork)
		{
			std::vector<C> a(Ny), s(Ny);
#pragma omp for collapse(2)
			for (int k = 0; k < nz; k++) for (int i = 0; i < Nc; i++) {
				for (int c = 0; c < 3; c++) {
					for (int j = 0; j < Ny; j++) a[j] = F[c][fidx(i, j, k)];
					fy.run(a.data(), s.data(), +1);
					for (int j = 0; j < ny; j++) F[c][fidx(i, j, k)] = a[j];
				}
			}
		}
		// inverse x (C2R), scale, write
		const R scale = (R)1 / (R)((double)Nx * Ny * Nz);
#pragma omp parallel if((size_t)Nx * Ny * Nz >= kOmpMinWork)
		{
			std::vector<C> a(Nx), s(Nx);
#pragma omp for collapse(2)
			for (int k = 0; k < nz; k++) for (int j = 0; j < ny; j++) {
				for (int c = 0; c < 3; c++) {
					for (int i = 0; i < Nc; i++) a[i] = F[c][fidx(i, j, k)];
					for (int i = Nc; i < Nx; i++) a[i] = std::conj(a[Nx - i]);
					a[0] = C(a[0].real(), 0); a[Nx / 2] = C(a[Nx / 2].real(), 0);
					fx.run(a.data(), s.data(), +1);
					for (int i = 0; i < nx; i++) {
						size_t idx = ((size_t)i + (size_t)j * nx + (size_t)k * nx * ny) * 3 + c;
						R v = a[i].real() * scale;
						H[idx] = clear_out ? v : H[idx] + v;
					}
				}
			}
		}
	}
};

// ---------------------------------------------------------------------------------------
// Material parameters of one ferromagnetic mesh (per-cell Ms map optional).
struct Params {
	double Ms = 8e5, A = 1.3e-11, alpha = 0.02, K1 = 0.0, K2 = 0.0;
	double ea[3] = { 1, 0, 0 };
	double grel = 1.0;
	// applied field H(t) = Ha + amp * dir * sinc(w (t - t0)) + slope * t   (A/m)
	double Ha[3] = { 0, 0, 0 };
	int wave = 0;                      // 0 constant, 1 sinc pulse, 2 linear ramp
	double amp = 0, w = 0, t0 = 0, dir[3] = { 0, 0, 0 }, slope[3] = { 0, 0, 0 };
	bool demag = true, exchange = true, anisotropy = false, zeeman = true;
};

inline void applied_field(const Params& p, double t, double H[3])
{
	for (int c = 0; c < 3; c++) H[c] = p.Ha[c];
	if (p.wave == 1) {
		double x = p.w * (t - p.t0);
		double s = (std::fabs(x) < 1e-300) ? 1.0 : std::sin(x) / x;   // F_sinc, BorisLib/TEquation_Function.h:247
		for (int c = 0; c < 3; c++) H[c] += p.amp * p.dir[c] * s;
	}
	else if (p.wave == 2) {
		for (int c = 0; c < 3; c++) H[c] += p.slope[c] * t;
	}
}

// Demag evaluation speedup, slot that the next computed field replaces once all p slots
// are filled: the oldest, found from the order of the save times tdm
// (Boris/DemagCUDA.cpp:176-324).
inline int speedup_next_slot(int p, const double* tdm)
{
	if (p == 1) return 0;
	bool rising = true;
	for (int s = 1; s < p; s++) rising = rising && tdm[s] > tdm[s - 1];
	if (rising) return 0;
	for (int s = 0; s + 2 < p; s++) if (tdm[s] > tdm[s + 1]) return s + 1;
	return p - 1;
}

// Lagrange coefficients of the p saved fields at time t (Boris/DemagCUDA.cpp:365-467):
// numerator and denominator products in ascending slot order, in FP64; slots >= p get 0.
inline void speedup_lagrange(int p, const double* tdm, double t, double a[6])
{
	for (int s = 0; s < 6; s++) a[s] = 0;
	if (p == 1) { a[0] = 1; return; }
	for (int s = 0; s < p; s++) {
		double num = 1, den = 1;
		for (int q = 0; q < p; q++) if (q != s) { num *= (t - tdm[q]); den *= (tdm[s] - tdm[q]); }
		a[s] = num / den;
	}
}

// Evaluation methods (values of the reference's EVAL_ enum, Boris/DiffEq_Defs.h).
enum { EV_TEULER = 1, EV_RK4 = 2, EV_AHEUN = 7, EV_RKDP54 = 9 };

// One ferromagnetic mesh with its fields and ODE state: the reference's
// SuperMesh::AdvanceTime loop (Boris/SuperMeshSimulation.cpp:84-118) driving
// ODECommon_Base::Iterate (Boris/DiffEq_CommonBase_Iterate.cpp:63-1533).
template <typename R>
struct Sim {
	typedef R V3[3];
	int nx, ny, nz; size_t n;
	double h[3];
	Params p;
	std::vector<R> Msmap;              // optional per-cell Ms (empty = uniform p.Ms)
	std::vector<R> M, Heff, sM1, E[6];
	Demag<R>* demag = nullptr;
	// Layer stack: the nz planes are separate 2D layers (one ferromagnetic mesh each, as the
	// reference's SDemag stack, Boris/SDemag.cpp:173-300): demag from ext_demag (H = op(M)),
	// exchange within each layer, per-layer A, alpha, K1 (empty = p's value).
	std::function<void(const R*, R*)> ext_demag;
	bool layered = false;
	std::vector<double> layA, layAlpha, layK1;
	// stacks of 3D meshes: mesh index of each layer (empty: every layer its own mesh);
	// exchange couples consecutive layers of one mesh across z (spacing = layHz), and layHa
	// (3 per layer, empty = p's field) gives each layer a constant applied field
	std::vector<int> layGrp;
	std::vector<double> layHa;

	int method;
	double time = 0, stagetime = 0, dT, dT_last, dT_min = 1e-15, dT_max = 3e-12, dT_increase = 2, err_high_fail = 1e-4, lte = 0;
	int order, evalStep = 0, iteration = 0, stageiteration = 0;
	bool primed = false, available = true, renormalize = true;
	long long rejections = 0, evaluations = 0, convolutions = 0;

	// Demag evaluation speedup (DemagCUDA::UpdateField, Boris/DemagCUDA.cpp:106-468):
	// speedup = EVALSPEEDUP_ value (0 none, 1 step, 2 linear, 3 quadratic, 4 cubic,
	// 5 quartic, 6 quintic, Boris/DiffEq_Defs.h:121-125) = number of saved demag fields.
	// Saved fields hold the convolution minus the self-demag term; extrapolated
	// evaluations add the Lagrange combination of the saved fields and the self term
	// of the current M.
	int speedup = 0, num_saved = 0;
	double tdm[6] = { 0, 0, 0, 0, 0, 0 };
	std::vector<R> Hs[6];
	R selfc[3] = { 0, 0, 0 };
	std::vector<R> layselfc;           // layer stacks: 3 per layer
	std::vector<double> layHz;         // layer stacks: each layer's cell thickness
	bool link_dTspeedup = true;
	double dTspeedup = 0, time_speedup = 0;
	// periodic boundary conditions (Demag::Set_PBC + Mesh::Set_Magnetic_PBC,
	// Boris/Demag.cpp:137-148, Boris/MeshControl.cpp:49): images per axis for the demag
	// tensor, and wrap-around neighbours for exchange along every axis with images
	int pbc[3] = { 0, 0, 0 };
	int Nz_open = 0;

	Sim(int method_, int nx_, int ny_, int nz_, double hx, double hy, double hz, const Params& p_, int Nz_override = 0)
		: nx(nx_), ny(ny_), nz(nz_), p(p_), method(method_)
	{
		n = (size_t)nx * ny * nz; h[0] = hx; h[1] = hy; h[2] = hz;
		M.assign(n * 3, (R)0); Heff.assign(n * 3, (R)0); sM1.assign(n * 3, (R)0);
		for (int s = 0; s < 6; s++) E[s].assign(n * 3, (R)0);
		for (size_t i = 0; i < n; i++) M[i * 3] = (R)-p.Ms;   // default M = (-Ms, 0, 0)
		if (p.demag) demag = new Demag<R>(nx, ny, nz, hx, hy, hz, Nz_override);
		Nz_open = Nz_override;
		// method defaults, Boris/DiffEq_CommonBase_Control.cpp:10-140, Boris/DiffEq_Defs.h
		switch (method) {
		case EV_TEULER: dT = 0.5e-13; order = 2; break;
		case EV_AHEUN: dT = 0.5e-13; err_high_fail = 1e-4; dT_increase = 2; dT_max = 1e-12; dT_min = 1e-15; order = 2; break;
		case EV_RK4: dT = 0.5e-12; order = 4; break;
		case EV_RKDP54: dT = 0.5e-12; err_high_fail = 1e-4; dT_increase = 2; dT_max = 3e-12; dT_min = 1e-15; order = 5; break;
		default: throw std::runtime_error("unsupported evaluation method");
		}
		dT_last = dT;
	}
	~Sim() { delete demag; }

	R Ms_at(size_t i) const { return Msmap.empty() ? (R)p.Ms : Msmap[i]; }
	size_t layer_of(size_t i) const { return i / ((size_t)nx * ny); }
	R A_at(size_t i) const { return layA.empty() ? (R)p.A : (R)layA[layer_of(i)]; }
	R alpha_at(size_t i) const { return layAlpha.empty() ? (R)p.alpha : (R)layAlpha[layer_of(i)]; }
	R K1_at(size_t i) const { return layK1.empty() ? (R)p.K1 : (R)layK1[layer_of(i)]; }

	// Zeeman (sets), Demag, Exchange, Anisotropy (add) — Boris/MeshModules.cpp:131-135.
	void update_fields()
	{
		double Hst[3]; applied_field(p, stagetime, Hst);
		R Ha[3] = { (R)Hst[0], (R)Hst[1], (R)Hst[2] };
		for (size_t i = 0; i < n; i++) for (int c = 0; c < 3; c++)
			Heff[i * 3 + c] = !p.zeeman ? (R)0 : layHa.empty() ? Ha[c] : (R)layHa[layer_of(i) * 3 + c];
		if (ext_demag || demag) demag_update();
		if (p.exchange) add_exchange();
		if (p.anisotropy) add_anisotropy();
		evaluations++;
	}

	// EVALSPEEDUP_ mode and, for dTs > 0, an independent speedup time step
	// (ODECommon_Base::SetEvaluationSpeedup / SetSpeedupTimeStep,
	// Boris/DiffEq_CommonBase.h:252, Boris/DiffEq_CommonBase_Control.cpp:243-256).
	// Saved fields are dropped, as UpdateConfiguration does (Boris/DemagCUDA.cpp:76-103).
	void set_speedup(int mode, double dTs)
	{
		if (mode < 0 || mode > 6) throw std::runtime_error("bad evaluation speedup mode");
		speedup = mode; num_saved = 0;
		for (int s = 0; s < 6; s++) { Hs[s].clear(); tdm[s] = 0; }
		double sd[3];
		self_demag_pbc(h[0], h[1], h[2], nx, ny, nz, pbc[0], pbc[1], pbc[2], sd);   // SelfDemag_PBC
		for (int c = 0; c < 3; c++) selfc[c] = (R)sd[c];
		// a layer stack: each layer mesh's own coefficient (Boris/SDemagCUDA.cpp:583-1000)
		layselfc.clear();
		for (size_t l = 0; l < layHz.size(); l++) {
			self_demag(h[0], h[1], layHz[l], sd);
			for (int c = 0; c < 3; c++) layselfc.push_back((R)sd[c]);
		}
		if (dTs > 0) { link_dTspeedup = false; dTspeedup = dTs; time_speedup = time; }
		else { link_dTspeedup = true; dTspeedup = dT; }
	}

	// Periodic images per axis (the `pbc` command): the demag tensor is rebuilt with them and
	// exchange wraps around; saved speedup fields are dropped (UpdateConfiguration).
	void set_pbc(int px, int py, int pz)
	{
		if (layered) throw std::runtime_error("periodic boundaries on a layer stack");
		pbc[0] = px; pbc[1] = py; pbc[2] = pz;
		if (p.demag) {
			delete demag;
			demag = (px || py || pz) ? new Demag<R>(nx, ny, nz, h[0], h[1], h[2], 0, pbc)
			                         : new Demag<R>(nx, ny, nz, h[0], h[1], h[2], Nz_open);
		}
		if (speedup) set_speedup(speedup, link_dTspeedup ? 0.0 : dTspeedup);
	}

	// ODECommon_Base::Get_EvalStep_Time (Boris/DiffEq_CommonBase_Iterate.cpp:1536-1635):
	// the step start time plus dT times the tableau's c value of the stage about to run.
	double eval_step_time() const
	{
		static const double c_teuler[2] = { 0.0, 1.0 };
		static const double c_rk4[4] = { 0.0, 0.5, 0.5, 1.0 };
		static const double c_rkdp54[6] = { 0.0, 0.2, 0.3, 0.8, 8.0 / 9, 1.0 };
		switch (method) {
		case EV_TEULER: case EV_AHEUN: return time + dT * c_teuler[evalStep];
		case EV_RK4: return time + dT * c_rk4[evalStep];
		case EV_RKDP54: return time + dT * c_rkdp54[evalStep];
		}
		return time;
	}

	// ODECommon_Base::Check_Step_Update (Boris/DiffEq_CommonBase_Iterate.cpp:1639-1657)
	bool check_step_update()
	{
		if (!speedup) return true;
		if (!link_dTspeedup) {
			double te = eval_step_time();
			if (te < time_speedup + dTspeedup) return false;
			time_speedup = te;
			return true;
		}
		return available;
	}

	// self coefficient of component c of cell i
	R self_at(size_t i, int c) const { return layselfc.empty() ? selfc[c] : layselfc[layer_of(i) * 3 + c]; }

	// H = demag field of M (set): the mesh's convolution or the stack's multilayered one
	void convolve(const R* Min, R* H)
	{
		if (ext_demag) ext_demag(Min, H);
		else demag->convolute(Min, H, true);
		convolutions++;
	}

	// DemagCUDA::UpdateField with the evaluation speedup (Boris/DemagCUDA.cpp:106-468), and
	// SDemagCUDA::UpdateField's multilayered branch for a layer stack (Boris/SDemagCUDA.cpp:
	// 378-1000: the same saves and extrapolation per layer mesh, slot times shared).
	void demag_update()
	{
		const size_t n3 = n * 3;
		if (!speedup || (num_saved < speedup && !check_step_update())) {
			std::vector<R> Hd(n3);
			convolve(M.data(), Hd.data());
			for (size_t q = 0; q < n3; q++) Heff[q] += Hd[q];
			return;
		}
		if (check_step_update() || num_saved < speedup) {
			int slot;
			if (num_saved < speedup) { slot = num_saved; tdm[slot] = eval_step_time(); num_saved++; }
			else { slot = speedup_next_slot(speedup, tdm); if (speedup > 1) tdm[slot] = eval_step_time(); }
			std::vector<R>& S = Hs[slot];
			S.assign(n3, (R)0);
			convolve(M.data(), S.data());
			// Demag_EvalSpeedup_AddField_SubSelf (Boris/DemagCUDA.cu:44-57)
			for (size_t q = 0; q < n3; q++) { Heff[q] += S[q]; S[q] -= self_at(q / 3, (int)(q % 3)) * M[q]; }
		}
		else {
			// Demag_EvalSpeedup_AddExtrapField_AddSelf (Boris/DemagCUDA.cu:59-413)
			double ad[6]; speedup_lagrange(speedup, tdm, eval_step_time(), ad);
			R a[6]; for (int s = 0; s < 6; s++) a[s] = (R)ad[s];
			for (size_t q = 0; q < n3; q++) {
				R v = Hs[0][q] * a[0];
				for (int s = 1; s < speedup; s++) v += Hs[s][q] * a[s];
				v += self_at(q / 3, (int)(q % 3)) * M[q];
				Heff[q] += v;
			}
		}
	}

	// Heff += 2A/(mu0 Ms^2) * delsq_neu(M)  (Boris/Exchange.cpp:108-156,
	// BorisLib/VEC_VC_Del.h:10-130; free boundaries use the one-sided difference, periodic
	// ones the cell on the other side: (M[+] + M[-]) - 2M with the wrapped neighbour).
	void add_exchange()
	{
		int d[3] = { nx, ny, nz };
		size_t st[3] = { 1, (size_t)nx, (size_t)nx * ny };
#pragma omp parallel for collapse(2) if(n >= kOmpMinWork)
		for (int k = 0; k < nz; k++) for (int j = 0; j < ny; j++) for (int i = 0; i < nx; i++) {
			size_t idx = (size_t)i + (size_t)j * nx + (size_t)k * nx * ny;
			int ijk[3] = { i, j, k };
			R lap[3] = { 0, 0, 0 };
			const bool grouped = layered && !layGrp.empty();
			for (int a = 0; a < (layered && !grouped ? 2 : 3); a++) {
				R diff[3] = { 0, 0, 0 };
				bool lo = ijk[a] > 0, hi = ijk[a] < d[a] - 1;
				if (grouped && a == 2) {
					lo = lo && layGrp[k - 1] == layGrp[k];
					hi = hi && layGrp[k + 1] == layGrp[k];
				}
				for (int c = 0; c < 3; c++) {
					R m0 = M[idx * 3 + c];
					if (lo && hi) diff[c] = M[(idx + st[a]) * 3 + c] + M[(idx - st[a]) * 3 + c] - 2 * m0;
					else if (pbc[a] && d[a] > 1) {
						const size_t up = hi ? idx + st[a] : idx - (size_t)(d[a] - 1) * st[a];
						const size_t dn = lo ? idx - st[a] : idx + (size_t)(d[a] - 1) * st[a];
						diff[c] = M[up * 3 + c] + M[dn * 3 + c] - 2 * m0;
					}
					else if (hi) diff[c] = M[(idx + st[a]) * 3 + c] - m0;
					else if (lo) diff[c] = M[(idx - st[a]) * 3 + c] - m0;
					const R ha = (grouped && a == 2) ? (R)layHz[k] : (R)h[a];
					diff[c] /= (ha * ha);
				}
				for (int c = 0; c < 3; c++) lap[c] += diff[c];
			}
			R Ms = Ms_at(idx);
			R pre = (R)2 * A_at(idx) / ((R)kMU0 * Ms * Ms);
			for (int c = 0; c < 3; c++) Heff[idx * 3 + c] += pre * lap[c];
		}
	}

	// Heff += 2/(mu0 Ms) (m.e) (K1 + 2 K2 (1 - (m.e)^2)) e  (Boris/Anisotropy.cpp:82-115)
	void add_anisotropy()
	{
		R e[3] = { (R)p.ea[0], (R)p.ea[1], (R)p.ea[2] };
#pragma omp parallel for if(n >= kOmpMinWork)
		for (long long ii = 0; ii < (long long)n; ii++) {
			size_t i = (size_t)ii;
			R Ms = Ms_at(i);
			R dot = (M[i * 3] * e[0] + M[i * 3 + 1] * e[1] + M[i * 3 + 2] * e[2]) / Ms;
			R pre = ((R)2 / ((R)kMU0 * Ms)) * dot * (K1_at(i) + 2 * (R)p.K2 * (1 - dot * dot));
			for (int c = 0; c < 3; c++) Heff[i * 3 + c] += pre * e[c];
		}
	}

	// LLG right-hand side (Boris/DiffEqFM_Equations.cpp:12-24)
	void llg(size_t i, R out[3]) const
	{
		R Ms = Ms_at(i);
		R a = alpha_at(i);
		const R* m = &M[i * 3]; const R* H = &Heff[i * 3];
		R mxh[3] = { m[1] * H[2] - m[2] * H[1], m[2] * H[0] - m[0] * H[2], m[0] * H[1] - m[1] * H[0] };
		R mu[3] = { m[0] / Ms, m[1] / Ms, m[2] / Ms };
		R mmh[3] = { mu[1] * mxh[2] - mu[2] * mxh[1], mu[2] * mxh[0] - mu[0] * mxh[2], mu[0] * mxh[1] - mu[1] * mxh[0] };
		R pre = (R)(-kGAMMA * p.grel) / (1 + a * a);
		for (int c = 0; c < 3; c++) out[c] = pre * (mxh[c] + a * mmh[c]);
	}

	void renorm(size_t i)
	{
		R* m = &M[i * 3];
		R nrm = std::sqrt(m[0] * m[0] + m[1] * m[1] + m[2] * m[2]);
		if (nrm > 0) { R s = Ms_at(i) / nrm; m[0] *= s; m[1] *= s; m[2] *= s; }
	}

	static R mag3(const R v[3]) { return std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]); }

	// ODECommon_Base::SetAdaptiveTimeStep (Boris/DiffEq_CommonBase_Iterate.cpp:15-59)
	bool set_adaptive_time_step()
	{
		if (dT_min == dT_max) return true;
		if (lte > err_high_fail && dT > dT_min) {
			time -= dT; stagetime -= dT; iteration--; stageiteration--;
			double c = std::pow(err_high_fail * 0.8 / lte, 1.0 / (order + 1));
			if (c < 0.01) c = 0.01;
			dT *= c;
			if (dT < dT_min) dT = dT_min;
			available = false;
			rejections++;
			return false;
		}
		else if (lte > 0) {
			double c = std::pow(err_high_fail * 0.8 / lte, 1.0 / (order + 1));
			if (c > dT_increase) c = dT_increase;
			if (c < 0.01) c = 0.01;
			dT *= c;
			if (dT < dT_min) dT = dT_min;
			if (dT > dT_max) dT = dT_max;
		}
		return true;
	}

	void restore() { M = sM1; }

	void end_step() { evalStep = 0; time += dT; stagetime += dT; iteration++; stageiteration++; available = true; }

	// One Iterate() call: one stage of the chosen method.
	void iterate_stage()
	{
		dT_last = dT;
		const R dt = (R)dT;
		switch (method) {
		case EV_RKDP54: {
			if (evalStep == 0) {
				double lmax = 0;
#pragma omp parallel for reduction(max:lmax) if(n >= kOmpMinWork)
				for (long long ii = 0; ii < (long long)n; ii++) {
					size_t i = (size_t)ii; R rhs[3]; llg(i, rhs);
					R pred[3], d[3];
					for (int c = 0; c < 3; c++) {
						size_t q = i * 3 + c;
						pred[c] = sM1[q] + ((R)5179 * E[0][q] / (R)57600 + (R)7571 * E[2][q] / (R)16695 + (R)393 * E[3][q] / (R)640
							- (R)92097 * E[4][q] / (R)339200 + (R)187 * E[5][q] / (R)2100 + rhs[c] / (R)40) * dt;
						d[c] = M[q] - pred[c];
					}
					double l = (double)(mag3(d) / mag3(&M[i * 3]));
					lmax = std::max(lmax, l);
					for (int c = 0; c < 3; c++) E[0][i * 3 + c] = rhs[c];
				}
				available = false;
				if (primed) {
					dT_last = dT; lte = lmax;
					if (!set_adaptive_time_step()) { restore(); primed = false; break; }
				}
				else primed = true;
				const R dtn = (R)dT;
				for (size_t q = 0; q < n * 3; q++) { sM1[q] = M[q]; M[q] += E[0][q] * (dtn / 5); }
				evalStep++;
			}
			else if (evalStep <= 4) {
				int s = evalStep;
#pragma omp parallel for if(n >= kOmpMinWork)
				for (long long ii = 0; ii < (long long)n; ii++) {
					size_t i = (size_t)ii; R rhs[3]; llg(i, rhs);
					for (int c = 0; c < 3; c++) {
						size_t q = i * 3 + c;
						E[s][q] = rhs[c];
						R v;
						if (s == 1) v = ((R)3 * E[0][q] / (R)40 + (R)9 * E[1][q] / (R)40);
						else if (s == 2) v = ((R)44 * E[0][q] / (R)45 - (R)56 * E[1][q] / (R)15 + (R)32 * E[2][q] / (R)9);
						else if (s == 3) v = ((R)19372 * E[0][q] / (R)6561 - (R)25360 * E[1][q] / (R)2187 + (R)64448 * E[2][q] / (R)6561 - (R)212 * E[3][q] / (R)729);
						else v = ((R)9017 * E[0][q] / (R)3168 - (R)355 * E[1][q] / (R)33 + (R)46732 * E[2][q] / (R)5247 + (R)49 * E[3][q] / (R)176 - (R)5103 * E[4][q] / (R)18656);
						M[q] = sM1[q] + v * dt;
					}
				}
				evalStep++;
			}
			else {
#pragma omp parallel for if(n >= kOmpMinWork)
				for (long long ii = 0; ii < (long long)n; ii++) {
					size_t i = (size_t)ii; R rhs[3]; llg(i, rhs);
					for (int c = 0; c < 3; c++) {
						size_t q = i * 3 + c;
						E[5][q] = rhs[c];
						M[q] = sM1[q] + ((R)35 * E[0][q] / (R)384 + (R)500 * E[2][q] / (R)1113 + (R)125 * E[3][q] / (R)192
							- (R)2187 * E[4][q] / (R)6784 + (R)11 * E[5][q] / (R)84) * dt;
					}
					if (renormalize) renorm(i);
				}
				end_step();
			}
		} break;
		case EV_RK4: {
			int s = evalStep;
#pragma omp parallel for if(n >= kOmpMinWork)
			for (long long ii = 0; ii < (long long)n; ii++) {
				size_t i = (size_t)ii; R rhs[3]; llg(i, rhs);
				for (int c = 0; c < 3; c++) {
					size_t q = i * 3 + c;
					if (s == 0) { sM1[q] = M[q]; E[0][q] = rhs[c]; M[q] += E[0][q] * (dt / 2); }
					else if (s == 1) { E[1][q] = rhs[c]; M[q] = sM1[q] + E[1][q] * (dt / 2); }
					else if (s == 2) { E[2][q] = rhs[c]; M[q] = sM1[q] + E[2][q] * dt; }
					else M[q] = sM1[q] + (E[0][q] + 2 * E[1][q] + 2 * E[2][q] + rhs[c]) * (dt / 6);
				}
				if (s == 3 && renormalize) renorm(i);
			}
			if (s < 3) { evalStep++; available = false; }
			else end_step();
		} break;
		case EV_TEULER:
		case EV_AHEUN: {
			if (evalStep == 0) {
#pragma omp parallel for if(n >= kOmpMinWork)
				for (long long ii = 0; ii < (long long)n; ii++) {
					size_t i = (size_t)ii; R rhs[3]; llg(i, rhs);
					for (int c = 0; c < 3; c++) { size_t q = i * 3 + c; sM1[q] = M[q]; M[q] += rhs[c] * dt; }
				}
				evalStep = 1; available = false;
			}
			else {
				double lmax = 0;
#pragma omp parallel for reduction(max:lmax) if(n >= kOmpMinWork)
				for (long long ii = 0; ii < (long long)n; ii++) {
					size_t i = (size_t)ii; R rhs[3]; llg(i, rhs);
					R saved[3] = { M[i * 3], M[i * 3 + 1], M[i * 3 + 2] };
					for (int c = 0; c < 3; c++) { size_t q = i * 3 + c; M[q] = (sM1[q] + M[q] + rhs[c] * dt) / 2; }
					if (renormalize) renorm(i);
					if (method == EV_AHEUN) {
						R d[3] = { M[i * 3] - saved[0], M[i * 3 + 1] - saved[1], M[i * 3 + 2] - saved[2] };
						lmax = std::max(lmax, (double)(mag3(d) / mag3(&M[i * 3])));
					}
				}
				end_step();
				if (method == EV_AHEUN) {
					dT_last = dT; lte = lmax;
					if (!set_adaptive_time_step()) restore();
				}
			}
		} break;
		}
	}

	// SuperMesh::AdvanceTime: fields then one stage, until the step is solved.
	void advance_time()
	{
		do { update_fields(); iterate_stage(); } while (!available);
	}

	void average_m(double out[3]) const
	{
		double s[3] = { 0, 0, 0 };
		for (size_t i = 0; i < n; i++) { double ms = (double)Ms_at(i); for (int c = 0; c < 3; c++) s[c] += (double)M[i * 3 + c] / ms; }
		for (int c = 0; c < 3; c++) out[c] = s[c] / (double)n;
	}
};

// ---------------------------------------------------------------------------------------
// BorisRand LCG (BorisLib/BLib_prng.h:18-90, single-thread sequence) and
// FMesh::SetRandomMag (Boris/Mesh_Ferromagnetic_Control.cpp:230-255).
struct BorisRand {
	uint32_t x;
	explicit BorisRand(uint32_t seed) : x(seed) {}
	double rand() { x = 1664525u * x + 1013904223u; return (double)x / (uint32_t)4294967295u; }
};

template <typename R>
inline void set_random_mag(std::vector<R>& M, size_t n, uint32_t seed)
{
	BorisRand g(seed);
	for (size_t i = 0; i < n; i++) {
		double ms = std::sqrt((double)M[i * 3] * M[i * 3] + (double)M[i * 3 + 1] * M[i * 3 + 1] + (double)M[i * 3 + 2] * M[i * 3 + 2]);
		double th = std::acos(1 - 2 * g.rand());
		double ph = g.rand() * 2 * kPI;
		M[i * 3 + 0] = (R)(ms * std::cos(ph) * std::sin(th));
		M[i * 3 + 1] = (R)(ms * std::sin(ph) * std::sin(th));
		M[i * 3 + 2] = (R)(ms * std::cos(th));
	}
}

} // namespace oracle
