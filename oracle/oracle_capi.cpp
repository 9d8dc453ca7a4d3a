// ORACLE — TEST INFRASTRUCTURE ONLY.  C-ABI over the oracle for the Python tests
// (ctypes) and bench.py's cpu_baseline leg.  Nothing in the product links it.
#include "oracle_core.hpp"
#include <memory>
#include <type_traits>
#include "oracle_mlayer.hpp"
#include "oracle_adjoint.hpp"

#include <cstring>
#include <exception>

using namespace oracle;

extern "C" {

// --- tensors and kernels ---------------------------------------------------------------

// Full-N tensors, -N convention.  Ddiag, Dodiag: Nx*Ny*Nz*3 doubles.
int or_tensor3d(int nx, int ny, int nz, int Nx, int Ny, int Nz, double hx, double hy, double hz,
	double* Ddiag, double* Dodiag)
{
	try {
		double hm = std::max(hx, std::max(hy, hz));
		std::vector<double> Dd, Do;
		tensor3d(nx, ny, nz, Nx, Ny, Nz, hx / hm, hy / hm, hz / hm, Dd, Do);
		std::memcpy(Ddiag, Dd.data(), Dd.size() * sizeof(double));
		std::memcpy(Dodiag, Do.data(), Do.size() * sizeof(double));
		return 0;
	}
	catch (...) { return 1; }
}

// Ddiag: Nx*Ny*3, Dxy: Nx*Ny
int or_tensor2d(int nx, int ny, int Nx, int Ny, double hx, double hy, double hz, double* Ddiag, double* Dxy)
{
	try {
		double hm = std::max(hx, std::max(hy, hz));
		std::vector<double> Dd, Do;
		tensor2d(nx, ny, Nx, Ny, hx / hm, hy / hm, hz / hm, Dd, Do);
		std::memcpy(Ddiag, Dd.data(), Dd.size() * sizeof(double));
		std::memcpy(Dxy, Do.data(), Do.size() * sizeof(double));
		return 0;
	}
	catch (...) { return 1; }
}

// 2D z-shifted tensor between source cell s and destination cell d (already normalised),
// shift shz: Ddiag and Dodiag Nx*Ny*3 each (-N convention).
int or_tensor2d_zshift(int nx, int ny, int Nx, int Ny, const double* s3, const double* d3, double shz, double* Ddiag, double* Dodiag)
{
	try {
		std::vector<double> Dd, Do;
		tensor2d_zshift(nx, ny, Nx, Ny, s3, d3, shz, Dd, Do);
		std::memcpy(Ddiag, Dd.data(), Dd.size() * sizeof(double));
		std::memcpy(Dodiag, Do.data(), Do.size() * sizeof(double));
		return 0;
	}
	catch (...) { return 1; }
}

// Real kernels of a z-shifted 2D tensor: Kd, Ko (Ny/2+1)*(Nx/2+1)*3 each.
// 2D tensor shifted in-plane (sh3[0] or sh3[1] nonzero) and along z: Ddiag and Dodiag Nx*Ny*3 each
int or_tensor2d_shifted(int nx, int ny, int Nx, int Ny, const double* s3, const double* d3, const double* sh3, double* Ddiag, double* Dodiag)
{
	try {
		std::vector<double> Dd, Do;
		tensor2d_shifted(nx, ny, Nx, Ny, s3, d3, sh3, Dd, Do);
		std::memcpy(Ddiag, Dd.data(), Dd.size() * sizeof(double));
		std::memcpy(Dodiag, Do.data(), Do.size() * sizeof(double));
		return 0;
	}
	catch (...) { return 1; }
}

// complex kernels of such a tensor: Nc * Ny * 6 complex (re, im interleaved), kx fastest
int or_kernels2d_complex(int Nx, int Ny, const double* Ddiag, const double* Dodiag, double* K)
{
	try {
		const size_t NN = (size_t)Nx * Ny * 3;
		std::vector<double> Dd(Ddiag, Ddiag + NN), Do(Dodiag, Dodiag + NN);
		std::vector<std::complex<double>> k;
		kernels2d_complex(Nx, Ny, Dd, Do, k);
		std::memcpy(K, k.data(), k.size() * sizeof(std::complex<double>));
		return 0;
	}
	catch (...) { return 1; }
}

// the same with px, py periodic images along x and y (N = n along a periodic axis)
int or_tensor2d_zshift_pbc(int Nx, int Ny, const double* s3, const double* d3, double shz, int px, int py, double* Ddiag, double* Dodiag)
{
	try {
		std::vector<double> Dd, Do;
		tensor2d_zshift_pbc(Nx, Ny, s3, d3, shz, px, py, Dd, Do);
		std::memcpy(Ddiag, Dd.data(), Dd.size() * sizeof(double));
		std::memcpy(Dodiag, Do.data(), Do.size() * sizeof(double));
		return 0;
	}
	catch (...) { return 1; }
}

int or_kernels2d_zshift(int Nx, int Ny, const double* Ddiag, const double* Dodiag, double* Kd, double* Ko)
{
	try {
		const size_t NN = (size_t)Nx * Ny * 3;
		std::vector<double> Dd(Ddiag, Ddiag + NN), Do(Dodiag, Dodiag + NN), kd, ko;
		kernels2d_zshift(Nx, Ny, Dd, Do, kd, ko);
		std::memcpy(Kd, kd.data(), kd.size() * sizeof(double));
		std::memcpy(Ko, ko.data(), ko.size() * sizeof(double));
		return 0;
	}
	catch (...) { return 1; }
}

// Multilayered 2D demag of L stacked layers of nx*ny cells.
void* or_mlayer_create(int L, int nx, int ny, double hx, double hy, const double* hz, const double* z0)
{
	try { return new MLayer<double>(L, nx, ny, hx, hy, hz, z0); }
	catch (...) { return nullptr; }
}
// with px, py periodic images along x and y
void* or_mlayer_create_pbc(int L, int nx, int ny, double hx, double hy, const double* hz, const double* z0, int px, int py)
{
	try { return new MLayer<double>(L, nx, ny, hx, hy, hz, z0, px, py); }
	catch (...) { return nullptr; }
}
// layers whose grids are offset in-plane by rx[l], ry[l] (m)
void* or_mlayer_create_offset(int L, int nx, int ny, double hx, double hy, const double* hz, const double* z0, const double* rx, const double* ry)
{
	try { return new MLayer<double>(L, nx, ny, hx, hy, hz, z0, 0, 0, rx, ry); }
	catch (...) { return nullptr; }
}
void or_mlayer_destroy(void* h) { delete (MLayer<double>*)h; }
// M, H: L*nx*ny*3 doubles, layer-major
void or_mlayer_convolute(void* h, const double* M, double* H) { ((MLayer<double>*)h)->convolute(M, H); }
void or_mlayer_convolute_transposed(void* h, const double* M, double* H) { ((MLayer<double>*)h)->convolute(M, H, true); }

void or_self_demag(double hx, double hy, double hz, double* out3) { self_demag(hx, hy, hz, out3); }
void or_self_demag_pbc(double hx, double hy, double hz, int nx, int ny, int nz, int px, int py, int pz, double* out3)
{
	self_demag_pbc(hx, hy, hz, nx, ny, nz, px, py, pz, out3);
}
double or_newell_f(double x, double y, double z) { return nw_f(x, y, z); }
double or_newell_g(double x, double y, double z) { return nw_g(x, y, z); }
double or_asympt_diag(double hx, double hy, double hz, double x, double y, double z) { return AsymptDiag(hx, hy, hz).eval(x, y, z); }
double or_asympt_offdiag(double hx, double hy, double hz, double x, double y, double z) { return AsymptOffDiag(hx, hy, hz).eval(x, y, z); }

// Kernels: 3D Kdiag, Kodiag each (Nx/2+1)(Ny/2+1)(Nz/2+1)*3; 2D Kdiag (..)*3, Kodiag (..)*1
int or_kernels(int nx, int ny, int nz, int Nx, int Ny, int Nz, double hx, double hy, double hz, double* Kdiag, double* Kodiag)
{
	try {
		Kernels K = make_kernels(nx, ny, nz, Nx, Ny, Nz, hx, hy, hz);
		std::memcpy(Kdiag, K.Kdiag.data(), K.Kdiag.size() * sizeof(double));
		std::memcpy(Kodiag, K.Kodiag.data(), K.Kodiag.size() * sizeof(double));
		return 0;
	}
	catch (...) { return 1; }
}

// Periodic tensors (stored as -N, N = n along periodic axes): 3D Dd, Do each Nx*Ny*Nz*3;
// 2D (Nz = 1) Dd Nx*Ny*3 and Do Nx*Ny (Dxy only).  h normalised by the caller.
int or_tensor_pbc(int Nx, int Ny, int Nz, double hx, double hy, double hz, int px, int py, int pz, double* Dd, double* Do)
{
	try {
		std::vector<double> a, b;
		if (Nz == 1) tensor2d_pbc(Nx, Ny, hx, hy, hz, px, py, pz, a, b);
		else tensor3d_pbc(Nx, Ny, Nz, hx, hy, hz, px, py, pz, a, b);
		std::memcpy(Dd, a.data(), a.size() * sizeof(double));
		std::memcpy(Do, b.data(), b.size() * sizeof(double));
		return 0;
	}
	catch (...) { return 1; }
}

// Complex FFT of length n (interleaved re,im doubles), sign -1 / +1, unnormalised.
int or_fft(int n, double* data, int sign)
{
	Fft<double> f(n);
	std::vector<std::complex<double>> a(n), s(n);
	for (int i = 0; i < n; i++) a[i] = { data[2 * i], data[2 * i + 1] };
	f.run(a.data(), s.data(), sign);
	for (int i = 0; i < n; i++) { data[2 * i] = a[i].real(); data[2 * i + 1] = a[i].imag(); }
	return 0;
}

// --- demag convolution -----------------------------------------------------------------

void* or_demag_create(int nx, int ny, int nz, double hx, double hy, double hz, int Nz_override, int fp32)
{
	try {
		if (fp32) return new Demag<float>(nx, ny, nz, hx, hy, hz, Nz_override);
		return new Demag<double>(nx, ny, nz, hx, hy, hz, Nz_override);
	}
	catch (...) { return nullptr; }
}
// periodic boundary conditions: pbc[3] images per axis
void* or_demag_create_pbc(int nx, int ny, int nz, double hx, double hy, double hz, const int* pbc, int fp32)
{
	try {
		if (fp32) return new Demag<float>(nx, ny, nz, hx, hy, hz, 0, pbc);
		return new Demag<double>(nx, ny, nz, hx, hy, hz, 0, pbc);
	}
	catch (...) { return nullptr; }
}

void or_demag_destroy(void* h, int fp32)
{
	if (fp32) delete (Demag<float>*)h; else delete (Demag<double>*)h;
}

// M, H: n*3 doubles; in FP32 mode values are rounded to float on entry.
int or_demag_convolute(void* h, int fp32, const double* M, double* H, int clear_out)
{
	if (!fp32) { ((Demag<double>*)h)->convolute(M, H, clear_out != 0); return 0; }
	Demag<float>* d = (Demag<float>*)h;
	size_t n = (size_t)d->nx * d->ny * d->nz * 3;
	std::vector<float> m(M, M + n), hh(H, H + n);
	d->convolute(m.data(), hh.data(), clear_out != 0);
	for (size_t i = 0; i < n; i++) H[i] = hh[i];
	return 0;
}

// --- full simulation (fields + ODE) ----------------------------------------------------

struct OrSim { int fp32; void* s; };

// params: Ms, A, alpha, K1, K2, ea[3], Ha[3], wave, amp, w, t0, dir[3], slope[3] (22 doubles)
// flags: bit0 demag, bit1 exchange, bit2 anisotropy, bit3 zeeman
static Params unpack_params(const double* prm, int flags)
{
	Params p;
	p.Ms = prm[0]; p.A = prm[1]; p.alpha = prm[2]; p.K1 = prm[3]; p.K2 = prm[4];
	for (int c = 0; c < 3; c++) { p.ea[c] = prm[5 + c]; p.Ha[c] = prm[8 + c]; }
	p.wave = (int)prm[11]; p.amp = prm[12]; p.w = prm[13]; p.t0 = prm[14];
	for (int c = 0; c < 3; c++) { p.dir[c] = prm[15 + c]; p.slope[c] = prm[18 + c]; }
	p.demag = flags & 1; p.exchange = (flags >> 1) & 1; p.anisotropy = (flags >> 2) & 1; p.zeeman = (flags >> 3) & 1;
	return p;
}

// Discrete adjoint of the fixed-step RK4 rollout (FP64): per-cell Ms map Ms[n], h[3].
void* or_adj_create(int nx, int ny, int nz, const double* h, const double* prm, int flags, const double* Ms,
	double dt, int nsteps, int Nz_override)
{
	try { return new AdjRK4(nx, ny, nz, h, unpack_params(prm, flags), Ms, dt, nsteps, Nz_override); }
	catch (...) { return nullptr; }
}
void or_adj_destroy(void* h) { delete (AdjRK4*)h; }

// keep M every K steps only and recompute segments in the backward pass (0: whole tape)
void or_adj_set_checkpoint(void* h, int K) { ((AdjRK4*)h)->ckK = K; }
// periodic images per axis (demag tensor with images, wrapped exchange)
void or_adj_set_pbc(void* h, int px, int py, int pz) { ((AdjRK4*)h)->set_pbc(px, py, pz); }
// forward from M0 = Ms m0 (m0: n*3); final M into Mf; returns the loss against mt
double or_adj_forward(void* h, const double* m0, const double* mt, double* Mf)
{
	AdjRK4* a = (AdjRK4*)h;
	a->forward(m0, Mf);
	return a->loss(Mf, mt);
}
// gradients after or_adj_forward: dL/dMs (n) and gz = (dL/d amp, dL/d slope xyz)
void or_adj_backward(void* h, const double* m0, const double* mt, double* gMs, double* gz)
{
	((AdjRK4*)h)->backward(m0, mt, gMs, gz);
}
// layer stack for the adjoint (create it with demag off; with_demag adds the multilayered demag)
void or_adj_set_layers(void* h, int L, double hx, double hy, const double* hz, const double* z0, int with_demag,
	const double* A, const double* alpha, const double* K1)
{
	((AdjRK4*)h)->set_layers(L, hx, hy, hz, z0, with_demag != 0, A, alpha, K1);
}

void* or_sim_create(int method, int nx, int ny, int nz, double hx, double hy, double hz,
	const double* prm, int flags, int Nz_override, int fp32)
{
	try {
		Params p;
		p.Ms = prm[0]; p.A = prm[1]; p.alpha = prm[2]; p.K1 = prm[3]; p.K2 = prm[4];
		for (int c = 0; c < 3; c++) { p.ea[c] = prm[5 + c]; p.Ha[c] = prm[8 + c]; }
		p.wave = (int)prm[11]; p.amp = prm[12]; p.w = prm[13]; p.t0 = prm[14];
		for (int c = 0; c < 3; c++) { p.dir[c] = prm[15 + c]; p.slope[c] = prm[18 + c]; }
		p.demag = flags & 1; p.exchange = (flags >> 1) & 1; p.anisotropy = (flags >> 2) & 1; p.zeeman = (flags >> 3) & 1;
		OrSim* o = new OrSim;
		o->fp32 = fp32;
		if (fp32) o->s = new Sim<float>(method, nx, ny, nz, hx, hy, hz, p, Nz_override);
		else o->s = new Sim<double>(method, nx, ny, nz, hx, hy, hz, p, Nz_override);
		return o;
	}
	catch (...) { return nullptr; }
}

void or_sim_destroy(void* h)
{
	OrSim* o = (OrSim*)h;
	if (o->fp32) delete (Sim<float>*)o->s; else delete (Sim<double>*)o->s;
	delete o;
}

#define SIM_DISPATCH(o, body) do { if ((o)->fp32) { Sim<float>* s = (Sim<float>*)(o)->s; body; } else { Sim<double>* s = (Sim<double>*)(o)->s; body; } } while (0)

void or_sim_set_M(void* h, const double* M)
{
	OrSim* o = (OrSim*)h;
	SIM_DISPATCH(o, { for (size_t i = 0; i < s->n * 3; i++) s->M[i] = M[i]; });
}

void or_sim_get_M(void* h, double* M)
{
	OrSim* o = (OrSim*)h;
	SIM_DISPATCH(o, { for (size_t i = 0; i < s->n * 3; i++) M[i] = s->M[i]; });
}

void or_sim_get_Heff(void* h, double* H)
{
	OrSim* o = (OrSim*)h;
	SIM_DISPATCH(o, { for (size_t i = 0; i < s->n * 3; i++) H[i] = s->Heff[i]; });
}

// Layer stack (Sim::ext_demag): the sim's nz planes become L 2D layers (hx, hy, hz[l],
// bottom faces z0[l]) coupled by the multilayered demag; per-layer Ms, A, alpha, K1
// (NULL keeps the sim's value).  Create the sim with demag off; with_demag adds the
// multilayered demag, periodic along x and y when the sim's images were set first
// (or_sim_set_pbc before this call).
void or_sim_set_layers(void* h, int L, double hx, double hy, const double* hz, const double* z0, int with_demag,
	const double* Ms, const double* A, const double* alpha, const double* K1)
{
	OrSim* o = (OrSim*)h;
	SIM_DISPATCH(o, {
		s->layered = true;
		s->layHz.assign(hz, hz + L);
		if (with_demag) {
			typedef std::remove_reference<decltype(s->M[0])>::type RR;
			std::shared_ptr<MLayer<RR>> ml(new MLayer<RR>(L, s->nx, s->ny, hx, hy, hz, z0, s->pbc[0], s->pbc[1]));
			s->ext_demag = [ml](const RR* M, RR* H) { ml->convolute(M, H); };
		}
		if (A) s->layA.assign(A, A + L);
		if (alpha) s->layAlpha.assign(alpha, alpha + L);
		if (K1) s->layK1.assign(K1, K1 + L);
		if (Ms) {
			s->Msmap.resize(s->n);
			const size_t nl = (size_t)s->nx * s->ny;
			for (size_t i = 0; i < s->n; i++) s->Msmap[i] = Ms[i / nl];
		}
	});
}

void or_sim_set_Msmap(void* h, const double* Ms)
{
	OrSim* o = (OrSim*)h;
	SIM_DISPATCH(o, { if (!Ms) s->Msmap.clear(); else { s->Msmap.resize(s->n); for (size_t i = 0; i < s->n; i++) s->Msmap[i] = Ms[i]; } });
}

// dT0 (<=0 keeps the method default)
void or_sim_set_dt(void* h, double dT)
{
	OrSim* o = (OrSim*)h;
	SIM_DISPATCH(o, { s->dT = dT; s->dT_last = dT; });
}

// Step-control overrides: err_fail, dT_increase, dT_min, dT_max (<=0 keeps defaults).
void or_sim_set_astep(void* h, double err, double inc, double dtmin, double dtmax)
{
	OrSim* o = (OrSim*)h;
	SIM_DISPATCH(o, { if (err > 0) s->err_high_fail = err; if (inc > 0) s->dT_increase = inc; if (dtmin > 0) s->dT_min = dtmin; if (dtmax > 0) s->dT_max = dtmax; });
}

void or_sim_update_fields(void* h)
{
	OrSim* o = (OrSim*)h;
	SIM_DISPATCH(o, { s->update_fields(); });
}

void or_sim_iterate_stage(void* h)
{
	OrSim* o = (OrSim*)h;
	SIM_DISPATCH(o, { s->iterate_stage(); });
}

// Advance until the reference's iteration counter reaches `iterations`.
void or_sim_run_to_iteration(void* h, int iterations)
{
	OrSim* o = (OrSim*)h;
	SIM_DISPATCH(o, { while (s->iteration < iterations) s->advance_time(); });
}

void or_sim_advance_time(void* h)
{
	OrSim* o = (OrSim*)h;
	SIM_DISPATCH(o, { s->advance_time(); });
}

// state: time, dT, lte, iteration, evalStep, primed, rejections, evaluations, stagetime
void or_sim_state(void* h, double* st)
{
	OrSim* o = (OrSim*)h;
	SIM_DISPATCH(o, {
		st[0] = s->time; st[1] = s->dT; st[2] = s->lte; st[3] = s->iteration; st[4] = s->evalStep;
		st[5] = s->primed; st[6] = (double)s->rejections; st[7] = (double)s->evaluations; st[8] = s->stagetime; });
}

// Demag evaluation speedup: mode 0..6 (EVALSPEEDUP_), dTs > 0 unlinks the speedup step.
// stacks of 3D meshes: grp (L ints) mesh index per layer or null; Ha (3L) or null
void or_sim_set_layer_meshes(void* h, int L, const int* grp, const double* Ha)
{
	OrSim* o = (OrSim*)h;
	SIM_DISPATCH(o, {
		s->layGrp.clear(); s->layHa.clear();
		if (grp) s->layGrp.assign(grp, grp + L);
		if (Ha) s->layHa.assign(Ha, Ha + 3 * L);
	});
}

int or_sim_set_pbc(void* h, int px, int py, int pz)
{
	OrSim* o = (OrSim*)h;
	try { SIM_DISPATCH(o, { s->set_pbc(px, py, pz); }); return 0; }
	catch (...) { return 1; }
}

int or_sim_set_speedup(void* h, int mode, double dTs)
{
	OrSim* o = (OrSim*)h;
	try { SIM_DISPATCH(o, { s->set_speedup(mode, dTs); }); return 0; }
	catch (...) { return 1; }
}

// num_saved, tdm[6], convolutions (8 doubles)
void or_sim_speedup_info(void* h, double* out)
{
	OrSim* o = (OrSim*)h;
	SIM_DISPATCH(o, {
		out[0] = s->num_saved;
		for (int q = 0; q < 6; q++) out[1 + q] = s->tdm[q];
		out[7] = (double)s->convolutions; });
}

int or_speedup_next_slot(int p, const double* tdm) { return speedup_next_slot(p, tdm); }
void or_speedup_lagrange(int p, const double* tdm, double t, double* a6) { speedup_lagrange(p, tdm, t, a6); }

void or_sim_average_m(void* h, double* out3)
{
	OrSim* o = (OrSim*)h;
	SIM_DISPATCH(o, { s->average_m(out3); });
}

// Reference SetRandomMag restated (single-thread BorisRand sequence).
void or_set_random_mag(double* M, long long n, unsigned seed)
{
	std::vector<double> m(M, M + n * 3);
	set_random_mag(m, (size_t)n, seed);
	std::memcpy(M, m.data(), sizeof(double) * n * 3);
}

void or_borisrand(unsigned seed, int count, double* out)
{
	BorisRand g(seed);
	for (int i = 0; i < count; i++) out[i] = g.rand();
}

} // extern "C"
