// ORACLE — TEST INFRASTRUCTURE ONLY.
// CPU FP64 discrete adjoint of the fixed-step RK4 LLG rollout with demag, exchange, uniaxial
// anisotropy and a Zeeman field H(t) = Ha + amp * dir * sinc(w (t - t0)) or Ha + slope * t,
// per-cell Ms map; optionally a layer stack (the nz planes are 2D layers coupled by the
// multilayered demag, per-layer A, alpha, K1, exchange within layers).  Only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg may load or call anything under
// oracle/; the product (boris2_amd/) never does.
//
// The reference has no differentiation (a capability added beside its plain path); the
// forward scheme it differentiates is the oracle's own RK4 step (Boris/DiffEqFM_Evals_RK4.cpp:13-296
// with the field order of Boris/MeshModules.cpp:131-135, Zeeman at the step's start time,
// renormalisation to the cell's Ms after the step).
//
// Loss: L = 1 - (1/n) sum_i (M_i / Ms_i) . mt_i of the final state; initial state
// M0_i = Ms_i m0_i.  Gradients: dL/dMs_i (per cell), dL/d amp and dL/d slope.
// Parity: pinned by central finite differences of the same FP64 forward
// (tests/test_oracle_adjoint.py).
#pragma once

#include "oracle_core.hpp"
#include "oracle_mlayer.hpp"

#include <memory>

#include <cmath>
#include <vector>

namespace oracle {

struct AdjRK4 {
	int nx, ny, nz; size_t n;
	double h[3];
	Params p;
	std::vector<double> Ms;
	Demag<double>* dem = nullptr;
	// layer stack (set_layers): multilayered demag, per-layer A, alpha, K1 (empty: p's)
	std::shared_ptr<MLayer<double>> ml;
	bool layered = false;
	std::vector<double> layA, layAlpha, layK1;
	int pbc[3] = { 0, 0, 0 };   // periodic images per axis (set_pbc)
	double dt = 0; int nsteps = 0;
	// tape: per step the four stage inputs X1..X4 and the pre-renormalisation sum Y
	std::vector<std::vector<double>> X[4], Y;

	AdjRK4(int nx_, int ny_, int nz_, const double* hh, const Params& p_, const double* Ms_, double dt_, int nsteps_, int Nz_override)
		: nx(nx_), ny(ny_), nz(nz_), p(p_), dt(dt_), nsteps(nsteps_)
	{
		n = (size_t)nx * ny * nz;
		for (int a = 0; a < 3; a++) h[a] = hh[a];
		Ms.assign(Ms_, Ms_ + n);
		if (p.demag) dem = new Demag<double>(nx, ny, nz, h[0], h[1], h[2], Nz_override);
	}
	~AdjRK4() { delete dem; }

	// periodic images per axis (the `pbc` command, Boris/Demag.cpp:137-148,
	// Boris/MeshControl.cpp:49-65): the demag tensor sums the images and exchange wraps around
	void set_pbc(int px, int py, int pz)
	{
		pbc[0] = px; pbc[1] = py; pbc[2] = pz;
		if (dem) { delete dem; dem = new Demag<double>(nx, ny, nz, h[0], h[1], h[2], 0, pbc); }
	}

	void set_layers(int L, double hx, double hy, const double* hz, const double* z0, bool with_demag,
		const double* A, const double* alpha, const double* K1)
	{
		layered = true;
		// x, y images of set_pbc (called first) make the stack periodic in-plane
		if (with_demag) ml.reset(new MLayer<double>(L, nx, ny, hx, hy, hz, z0, pbc[0], pbc[1]));
		if (A) layA.assign(A, A + L);
		if (alpha) layAlpha.assign(alpha, alpha + L);
		if (K1) layK1.assign(K1, K1 + L);
	}
	size_t layer_of(size_t i) const { return i / ((size_t)nx * ny); }
	double A_at(size_t i) const { return layA.empty() ? p.A : layA[layer_of(i)]; }
	double alpha_at(size_t i) const { return layAlpha.empty() ? p.alpha : layAlpha[layer_of(i)]; }
	double K1_at(size_t i) const { return layK1.empty() ? p.K1 : layK1[layer_of(i)]; }

	double zeeman_scale(double t) const
	{
		if (p.wave != 1) return 0.0;
		const double x = p.w * (t - p.t0);
		return (std::fabs(x) < 1e-300) ? 1.0 : std::sin(x) / x;
	}

	// second differences / h^2 with free boundaries, or the wrapped neighbour at the ends of a
	// periodic axis (BorisLib/VEC_VC_Del.h:24-45) (the exchange Laplacian, unscaled)
	void laplacian(const double* V, double* out) const
	{
		const int d[3] = { nx, ny, nz };
		const size_t st[3] = { 1, (size_t)nx, (size_t)nx * ny };
#pragma omp parallel for collapse(2) if((size_t)nx * ny * nz >= kOmpMinWork)
		for (int k = 0; k < nz; k++) for (int j = 0; j < ny; j++) for (int i = 0; i < nx; i++) {
			const size_t idx = (size_t)i + (size_t)j * nx + (size_t)k * nx * ny;
			const int ijk[3] = { i, j, k };
			double l[3] = { 0, 0, 0 };
			for (int a = 0; a < (layered ? 2 : 3); a++) {
				const bool lo = ijk[a] > 0, hi = ijk[a] < d[a] - 1;
				for (int c = 0; c < 3; c++) {
					const double m0 = V[idx * 3 + c];
					double df = 0;
					if (lo && hi) df = V[(idx + st[a]) * 3 + c] + V[(idx - st[a]) * 3 + c] - 2 * m0;
					else if (pbc[a] && d[a] > 1) {
						const size_t up = hi ? idx + st[a] : idx - (size_t)(d[a] - 1) * st[a];
						const size_t dn = lo ? idx - st[a] : idx + (size_t)(d[a] - 1) * st[a];
						df = V[up * 3 + c] + V[dn * 3 + c] - 2 * m0;
					}
					else if (hi) df = V[(idx + st[a]) * 3 + c] - m0;
					else if (lo) df = V[(idx - st[a]) * 3 + c] - m0;
					l[c] += df / (h[a] * h[a]);
				}
			}
			for (int c = 0; c < 3; c++) out[idx * 3 + c] = l[c];
		}
	}

	double cex(size_t i) const { return 2 * A_at(i) / (kMU0 * Ms[i] * Ms[i]); }

	// Zeeman (Boris/Zeeman.cpp:164-300), demag, exchange (Boris/Exchange.cpp:108-156),
	// uniaxial anisotropy (Boris/Anisotropy.cpp:82-115); lap = the unscaled Laplacian
	void field(const double* Mv, double t, std::vector<double>& H, std::vector<double>& lap) const
	{
		double Hz[3]; applied_field(p, t, Hz);
		for (size_t i = 0; i < n; i++) for (int c = 0; c < 3; c++) H[i * 3 + c] = p.zeeman ? Hz[c] : 0.0;
		if (ml) {
			std::vector<double> Hd(3 * n);
			ml->convolute(Mv, Hd.data());
			for (size_t q = 0; q < 3 * n; q++) H[q] += Hd[q];
		}
		else if (dem) dem->convolute(Mv, H.data(), false);
		if (p.exchange) {
			laplacian(Mv, lap.data());
			for (size_t i = 0; i < n; i++) { const double ce = cex(i); for (int c = 0; c < 3; c++) H[i * 3 + c] += ce * lap[i * 3 + c]; }
		}
		if (p.anisotropy)
			for (size_t i = 0; i < n; i++) {
				const double d = dot(Mv + i * 3, p.ea) / Ms[i];
				const double pre = 2 / (kMU0 * Ms[i]) * d * (K1_at(i) + 2 * p.K2 * (1 - d * d));
				for (int c = 0; c < 3; c++) H[i * 3 + c] += pre * p.ea[c];
			}
	}

	static void cross(const double* a, const double* b, double* o) { o[0] = a[1] * b[2] - a[2] * b[1]; o[1] = a[2] * b[0] - a[0] * b[2]; o[2] = a[0] * b[1] - a[1] * b[0]; }
	static double dot(const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

	double pre(double al) const { return -kGAMMA * p.grel / (1 + al * al); }

	void rhs(const double* Mv, const double* H, double* k) const
	{
		for (size_t i = 0; i < n; i++) {
			const double al = alpha_at(i), pr = pre(al);
			const double* m = Mv + i * 3; const double* hh = H + i * 3;
			double mxh[3], mmh[3], mu[3] = { m[0] / Ms[i], m[1] / Ms[i], m[2] / Ms[i] };
			cross(m, hh, mxh); cross(mu, mxh, mmh);
			for (int c = 0; c < 3; c++) k[i * 3 + c] = pr * (mxh[c] + al * mmh[c]);
		}
	}

	// Checkpointing (0: keep the whole tape): with K > 0 the forward keeps M before every
	// K-th step only and the backward recomputes one segment's tape at a time from its
	// checkpoint, with the same arithmetic, so the gradients are the same bits.
	int ckK = 0;
	std::vector<std::vector<double>> ck;

	// steps st0 .. st1-1 from M (advanced in place); with tape, X and Y of step st go to
	// index st - st0 of the tape arrays
	void run_steps(std::vector<double>& M, int st0, int st1, bool tape)
	{
		std::vector<double> H(3 * n), lap(3 * n), k(3 * n), acc(3 * n);
		static const double cst[4] = { 0.5, 0.5, 1.0, 0.0 }, wgt[4] = { 1, 2, 2, 1 };
		for (int st = st0; st < st1; st++) {
			const double t = st * dt;
			std::vector<double> Xs = M;
			std::fill(acc.begin(), acc.end(), 0.0);
			for (int s = 0; s < 4; s++) {
				if (tape) X[s][st - st0] = Xs;
				field(Xs.data(), t, H, lap);
				rhs(Xs.data(), H.data(), k.data());
				for (size_t q = 0; q < 3 * n; q++) acc[q] += wgt[s] * k[q];
				if (s < 3) for (size_t q = 0; q < 3 * n; q++) Xs[q] = M[q] + cst[s] * dt * k[q];
			}
			std::vector<double> Yv(3 * n);
			for (size_t q = 0; q < 3 * n; q++) Yv[q] = M[q] + (dt / 6) * acc[q];
			for (size_t i = 0; i < n; i++) {
				const double nr = std::sqrt(dot(&Yv[i * 3], &Yv[i * 3]));
				for (int c = 0; c < 3; c++) M[i * 3 + c] = Ms[i] * Yv[i * 3 + c] / nr;
			}
			if (tape) Y[st - st0] = std::move(Yv);
		}
	}

	// forward rollout from M0 = Ms * m0; returns the final M, keeps the tape (or checkpoints)
	void forward(const double* m0, double* Mf)
	{
		std::vector<double> M(3 * n);
		for (size_t i = 0; i < n; i++) for (int c = 0; c < 3; c++) M[i * 3 + c] = Ms[i] * m0[i * 3 + c];
		if (ckK <= 0 || ckK >= nsteps) {
			for (int s = 0; s < 4; s++) X[s].assign(nsteps, {});
			Y.assign(nsteps, {});
			run_steps(M, 0, nsteps, true);
		}
		else {
			ck.clear();
			for (int st0 = 0; st0 < nsteps; st0 += ckK) {
				ck.push_back(M);
				run_steps(M, st0, std::min(nsteps, st0 + ckK), false);
			}
		}
		for (size_t q = 0; q < 3 * n; q++) Mf[q] = M[q];
	}

	double loss(const double* Mf, const double* mt) const
	{
		double s = 0;
		for (size_t i = 0; i < n; i++) s += dot(Mf + i * 3, mt + i * 3) / Ms[i];
		return 1.0 - s / (double)n;
	}

	// lamX = J_f(X)^T w (adjoint of k = f(X) at time t); accumulates dL/dMs and the Zeeman
	// gradients gz = (dL/d amp, dL/d slope).
	void vjp(const double* Xv, double t, const double* w, double* lamX, double* gMs, double* gz) const
	{
		std::vector<double> H(3 * n), lap(3 * n), u(3 * n), cu(3 * n), Nu(3 * n, 0.0), Lcu(3 * n, 0.0);
		field(Xv, t, H, lap);
		const double sc = zeeman_scale(t);
		double ga = 0, gs[3] = { 0, 0, 0 };
		for (size_t i = 0; i < n; i++) {
			const double al = alpha_at(i), pr = pre(al);
			const double* m = Xv + i * 3;
			const double* Hi = &H[i * 3]; const double* wi = w + i * 3;
			const double ms = Ms[i], mh = dot(m, Hi), mw = dot(m, wi), hw = dot(Hi, wi), mm = dot(m, m);
			double hxw[3], wxm[3], mxh[3], mmh[3];
			cross(Hi, wi, hxw); cross(wi, m, wxm); cross(m, Hi, mxh); cross(m, mxh, mmh);
			for (int c = 0; c < 3; c++) {
				lamX[i * 3 + c] = pr * (hxw[c] + (al / ms) * (wi[c] * mh + Hi[c] * mw - 2 * m[c] * hw));
				u[i * 3 + c] = pr * (wxm[c] + (al / ms) * (m[c] * mw - mm * wi[c]));
			}
			// Ms enters f through alpha/Ms and H through the exchange prefactor 2A/(mu0 Ms^2)
			// and the anisotropy field
			gMs[i] += -pr * al / (ms * ms) * dot(mmh, wi);
			if (p.exchange) gMs[i] += dot(&u[i * 3], &lap[i * 3]) * (-2 * cex(i) / ms);
			if (p.anisotropy) {
				// H_an = 2/mu0 s (K1 + 2K2 - 2K2 s^2/Ms^2)/Ms^2 e, s = m.e
				const double sm = dot(m, p.ea), d = sm / ms, ue = dot(&u[i * 3], p.ea), K1 = K1_at(i);
				const double dd = 2 / (kMU0 * ms * ms) * (K1 + 2 * p.K2 - 6 * p.K2 * d * d) * ue;
				for (int c = 0; c < 3; c++) lamX[i * 3 + c] += dd * p.ea[c];
				gMs[i] += 2 / kMU0 * sm * ue * (-2 * K1 - 4 * p.K2 + 8 * p.K2 * d * d) / (ms * ms * ms);
			}
			if (p.zeeman && p.wave == 1) ga += dot(&u[i * 3], p.dir) * sc;
			if (p.zeeman && p.wave == 2) for (int c = 0; c < 3; c++) gs[c] += u[i * 3 + c] * t;
			for (int c = 0; c < 3; c++) cu[i * 3 + c] = cex(i) * u[i * 3 + c];
		}
		gz[0] += ga;
		for (int c = 0; c < 3; c++) gz[1 + c] += gs[c];
		// through the fields: the demag operator is symmetric (N_ij = N_ji, N(r) = N(-r)), the
		// multilayered one has the exact transpose (N^T)_sd = N_ds^T, the free-boundary
		// Laplacian is symmetric; the exchange prefactor sits on the output side
		if (ml) ml->convolute(u.data(), Nu.data(), true);
		else if (dem) dem->convolute(u.data(), Nu.data(), true);
		if (p.exchange) laplacian(cu.data(), Lcu.data());
		for (size_t q = 0; q < 3 * n; q++) lamX[q] += Nu[q] + Lcu[q];
	}

	// gradients of L(forward(m0)) with the tape of the last forward(); gz = (amp, slope xyz)
	void backward(const double* m0, const double* mt, double* gMs, double* gz)
	{
		std::vector<double> lam(3 * n), lamY(3 * n), w(3 * n), lx(3 * n), lacc(3 * n);
		for (size_t i = 0; i < n; i++) gMs[i] = 0;
		for (int c = 0; c < 4; c++) gz[c] = 0;
		const bool seg = ckK > 0 && ckK < nsteps;
		// tape index of step st: st itself, or st - st0 of the segment recomputed last
		int base = 0;
		auto recompute = [&](int st) {
			const int sg = st / ckK, st0 = sg * ckK, st1 = std::min(nsteps, st0 + ckK);
			for (int s = 0; s < 4; s++) X[s].assign(st1 - st0, {});
			Y.assign(st1 - st0, {});
			std::vector<double> M = ck[sg];
			run_steps(M, st0, st1, true);
			base = st0;
		};
		if (seg) recompute(nsteps - 1);
		// final state M_N = Ms * Y/|Y| of the last step; dL/dM_N and the explicit Ms term
		{
			const std::vector<double>& Yv = Y[nsteps - 1 - base];
			for (size_t i = 0; i < n; i++) {
				const double nr = std::sqrt(dot(&Yv[i * 3], &Yv[i * 3]));
				double Mi[3];
				for (int c = 0; c < 3; c++) Mi[c] = Ms[i] * Yv[i * 3 + c] / nr;
				for (int c = 0; c < 3; c++) lam[i * 3 + c] = -mt[i * 3 + c] / ((double)n * Ms[i]);
				gMs[i] += dot(Mi, mt + i * 3) / ((double)n * Ms[i] * Ms[i]);
			}
		}
		static const double cst[4] = { 0.5, 0.5, 1.0, 0.0 }, wgt[4] = { 1, 2, 2, 1 };
		for (int st = nsteps - 1; st >= 0; st--) {
			const double t = st * dt;
			if (seg && st < base) recompute(st);
			const std::vector<double>& Yv = Y[st - base];
			// renormalisation M' = Ms Y/|Y|
			for (size_t i = 0; i < n; i++) {
				const double nr = std::sqrt(dot(&Yv[i * 3], &Yv[i * 3]));
				double yh[3] = { Yv[i * 3] / nr, Yv[i * 3 + 1] / nr, Yv[i * 3 + 2] / nr };
				const double yl = dot(yh, &lam[i * 3]);
				gMs[i] += yl;
				for (int c = 0; c < 3; c++) lamY[i * 3 + c] = (Ms[i] / nr) * (lam[i * 3 + c] - yh[c] * yl);
			}
			// Y = M + dt/6 (k1 + 2k2 + 2k3 + k4), X_{s+1} = M + c_s dt k_s
			for (size_t q = 0; q < 3 * n; q++) lacc[q] = lamY[q];      // adjoint of M_st
			std::vector<double> lamXnext(3 * n, 0.0);                  // adjoint of X_{s+1}
			for (int s = 3; s >= 0; s--) {
				for (size_t q = 0; q < 3 * n; q++) w[q] = (dt / 6) * wgt[s] * lamY[q] + (s < 3 ? cst[s] * dt * lamXnext[q] : 0.0);
				std::fill(lx.begin(), lx.end(), 0.0);
				vjp(X[s][st - base].data(), t, w.data(), lx.data(), gMs, gz);
				// X_s = M + c_{s-1} dt k_{s-1} (s > 0) or X_0 = M
				for (size_t q = 0; q < 3 * n; q++) lacc[q] += lx[q];
				lamXnext = lx;
				if (s == 0) break;
				// lx (adjoint of X_s) also feeds k_{s-1} through w in the next iteration
			}
			lam = lacc;
		}
		// M0 = Ms m0
		for (size_t i = 0; i < n; i++) gMs[i] += dot(m0 + i * 3, &lam[i * 3]);
	}
};

} // namespace oracle
