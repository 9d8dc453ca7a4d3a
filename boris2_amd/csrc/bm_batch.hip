// Batched fixed-step LLG integration of many independent meshes of uneven sizes.
//
// No reference counterpart: the reference integrates one mesh set per simulation.  Each
// mesh keeps its own demag convolution (kernels and spectra of a bm_demag handle), fields
// and parameters; what is shared is the launch: every pass of every field evaluation is
// one launch per distinct pass size (x passes per nx, y passes per N.y, z passes per N.z)
// with the meshes of that size as grid rows, and one stage launch covers all meshes.  A
// whole step (RK4: 4 evaluations, Heun: 2) is captured once into a hipGraph and replayed,
// so the host issues one graph launch per step however many meshes there are.
//
// Per-mesh arrays live concatenated (mesh b's cells at offset off[b]); the caller owns M.
// The per-mesh arguments (pointers, sizes, parameters, dT) sit in device tables read by the
// kernels, so changing dT rewrites a table and leaves the captured graph valid.
// Fixed-step methods (RK4, TEuler): time-dependent applied fields (sinc pulse, ramp) read the
// step time from a device scalar that the graph itself advances.
// RKDP54: every mesh keeps its own adaptive step (ODECommon_Base::Iterate's RKDP54 branch,
// Boris/DiffEq_CommonBase_Iterate.cpp:1214-1395, with SetAdaptiveTimeStep :15-59) in a
// device control block (BatchCtl).  One graph replay is one attempt of every active mesh:
// stage 0 (FSAL error estimate), a one-thread-per-mesh control kernel (accept with a new dT
// written into the stage tables, or reject: restore and re-prime), the advance, stages 1..5
// (skipped by rejected meshes) and the step end.  The host replays until every mesh has
// made its accepted steps, reading the control blocks back once per batch of replays.
//
// A mesh may be a layer stack (bm_layer_desc: its nz planes are 2D layers coupled by a
// multilayered demag, per-layer Ms, A, alpha, K1, as bm_ode_set_layers): its x and y passes
// join the launches of its sizes, its kernel product is one launch per layer count, and its
// stage arguments carry the per-layer values.
//
// A mesh may be periodic (bm_batch_set_pbc, the reference's `pbc` command): its demag handle
// sums the images, and along a periodic axis N = n, so its passes run the full-length
// variants; those form groups of their own (a group is a pass size and its periodicity),
// and its stage arguments wrap exchange around.
#include "bm_internal.h"
#include "bm_stage.h"

#include <algorithm>
#include <array>
#include <cstring>
#include <vector>
#include <string>
#include <cstdio>

using namespace bm;

namespace {

template <typename T> struct Tab {
	std::vector<T> h;
	T* d = nullptr;
};

// A pass-size group: its meshes and the launch shape of its pass(es) (x groups: [0] forward,
// [1] inverse; y groups: [0] forward, [1] inverse; others [0]), with the autotuned time.
struct Group {
	int size;
	bool full;                 // the pass runs along a periodic axis (N = n, every value data)
	std::vector<int> meshes;
	LaunchShape sh[2];
	double us[2] = { 0, 0 };
};

} // namespace

struct bm_batch_s {
	int nb = 0, mcode = 0, nstages = 0, use_graph = 1;
	std::vector<int> nx, ny, nz;
	std::vector<size_t> off, ncell;
	size_t ntot = 0, max_cells = 0;
	std::vector<bm_fm_params> p;
	std::vector<double> h;
	std::vector<std::array<int, 3>> pbc;   // periodic images per mesh and axis
	bool started = false;                  // the first bm_batch_run has begun
	std::vector<bm_demag_t> dm;
	std::vector<bm_mlayer_t> ml;   // layer stacks: their multilayered demag (else null)
	std::vector<std::vector<float>> lay;  // layer stacks: A, alpha, K1 per layer (3 x kMaxStackLayers)
	std::vector<int> lay_set;      // bits 0..2: A, alpha, K1 given per layer
	LayerTab* d_laytab = nullptr;  // per-mesh layer tables of the stacks (device)
	std::vector<int> lay_grouped, lay_field;   // stacks of 3D meshes: grouping and per-layer fields set
	float* Msmap = nullptr;        // per-cell Ms of the stacks with per-layer Ms (ntot floats)
	double* d_time = nullptr;      // step start time on the device (time-dependent fields)
	float* d_Ha = nullptr;         // applied field of each mesh at its step time (3 floats each)
	bool has_wave = false;         // some mesh has a time-dependent applied field
	double* d_dT = nullptr;
	float* M = nullptr;            // caller-owned, concatenated
	float* Mb = nullptr;
	float* sM1 = nullptr;
	float* E[6] = { nullptr, nullptr, nullptr, nullptr, nullptr, nullptr };
	// RKDP54: per-mesh step control (device) and its host copy (pinned, refreshed by runs)
	BatchCtl* d_ctl = nullptr;
	BatchCtl* h_ctl = nullptr;
	unsigned* d_lte = nullptr;
	float* Hd = nullptr;           // planar per mesh: 3 * ncell[b] at 3 * off[b]
	// pass groups and their tables
	std::vector<Group> gx, gy, gz, gy2, gml;
	std::vector<Group> gyz[2];             // thin meshes' fused y, z, y pass by N.y: [0] N.z 4, [1] N.z 8
	std::vector<Tab<FwdArgs>> fwd[2];      // [input parity][nx group]
	std::vector<Tab<PassBatch>> inv, ty, tz, ty2;
	std::vector<Tab<MlBatch>> tml;
	std::vector<Tab<PassBatch>> tyz[2];
	Tab<StageArgs> stage[2];               // [input parity]
	Tab<int> choff;                        // first flat chunk of kBS cells of each mesh (nb + 1)
	void* dev_tables = nullptr;
	size_t dev_bytes = 0;
	int ncu = 256;
	double dT = 0, time = 0;
	long long iteration = 0;
	hipStream_t gs = nullptr;
	hipEvent_t ev_in = nullptr, ev_out = nullptr;
	// side streams on which the pass-size groups of one pass run as parallel branches
	static constexpr int kMaxSide = 7;
	int nside = 0;
	hipStream_t side[kMaxSide] = {};
	hipEvent_t ev_fork = nullptr, ev_join[kMaxSide] = {};
	hipGraph_t graph = nullptr;
	hipGraphExec_t exec = nullptr;
	bool autotune = false, tune_radix = false, tuned = false;
};

namespace {

void free_all(bm_batch_s* o)
{
	if (o->exec) (void)hipGraphExecDestroy(o->exec);
	if (o->graph) (void)hipGraphDestroy(o->graph);
	if (o->ev_in) (void)hipEventDestroy(o->ev_in);
	if (o->ev_out) (void)hipEventDestroy(o->ev_out);
	if (o->gs) (void)hipStreamDestroy(o->gs);
	if (o->ev_fork) (void)hipEventDestroy(o->ev_fork);
	for (int k = 0; k < o->nside; k++) {
		if (o->ev_join[k]) (void)hipEventDestroy(o->ev_join[k]);
		if (o->side[k]) (void)hipStreamDestroy(o->side[k]);
	}
	for (bm_demag_t d : o->dm) if (d) bm_demag_destroy(d);
	for (bm_mlayer_t m : o->ml) if (m) bm_mlayer_destroy(m);
	if (o->Msmap) (void)hipFree(o->Msmap);
	if (o->d_time) (void)hipFree(o->d_time);
	if (o->d_Ha) (void)hipFree(o->d_Ha);
	if (o->d_laytab) (void)hipFree(o->d_laytab);
	if (o->d_dT) (void)hipFree(o->d_dT);
	if (o->Mb) (void)hipFree(o->Mb);
	if (o->sM1) (void)hipFree(o->sM1);
	for (int q = 0; q < 6; q++) if (o->E[q]) (void)hipFree(o->E[q]);
	if (o->d_ctl) (void)hipFree(o->d_ctl);
	if (o->h_ctl) (void)hipHostFree(o->h_ctl);
	if (o->d_lte) (void)hipFree(o->d_lte);
	if (o->Hd) (void)hipFree(o->Hd);
	if (o->dev_tables) (void)hipFree(o->dev_tables);
}

void add_to(std::vector<Group>& gs, int size, bool full, int b)
{
	for (Group& g : gs) if (g.size == size && g.full == full) { g.meshes.push_back(b); return; }
	gs.push_back(Group{ size, full, { b } });
}

StageArgs stage_args(bm_batch_s* o, int b, int parity)
{
	StageArgs a;
	std::memset(&a, 0, sizeof(a));
	const size_t off3 = 3 * o->off[b];
	float* Ma = o->M + off3;
	float* Mbb = o->Mb + off3;
	a.Min = parity ? Mbb : Ma;
	a.Mout = parity ? Ma : Mbb;
	a.Hd = o->Hd + off3;
	a.sM1 = o->sM1 + off3;
	for (int q = 0; q < 6; q++) a.E[q] = o->E[q] ? o->E[q] + off3 : nullptr;
	a.nx = o->nx[b]; a.ny = o->ny[b]; a.nz = o->nz[b];
	const double* h = &o->h[3 * b];
	a.ihx2 = (float)(1.0 / (h[0] * h[0])); a.ihy2 = (float)(1.0 / (h[1] * h[1])); a.ihz2 = (float)(1.0 / (h[2] * h[2]));
	const bm_fm_params& p = o->p[b];
	a.Ms = (float)p.Ms; a.A = (float)p.A; a.alpha = (float)p.alpha; a.K1 = (float)p.K1; a.K2 = (float)p.K2; a.grel = (float)p.grel;
	for (int c = 0; c < 3; c++) { a.ea[c] = (float)p.ea[c]; a.Ha[c] = (float)p.Ha[c]; }
	a.dT = (float)o->dT;
	a.modules = p.modules;
	a.per = (o->pbc[b][0] ? 1 : 0) | (o->pbc[b][1] ? 2 : 0) | (o->pbc[b][2] ? 4 : 0);
	a.map = 2;
	if (o->d_ctl) {
		a.ctl = o->d_ctl + b;
		a.lte = o->d_lte + b;
		a.dT = (float)o->h_ctl[b].dT;
	}
	if ((p.modules & BM_MOD_ZEEMAN) && (p.wave == 1 || p.wave == 2)) {
		a.tnow = o->d_ctl ? &o->d_ctl[b].time : o->d_time; a.wave = p.wave;
		a.Ha_dev = o->d_Ha + 3 * b;
		a.amp = p.amp; a.w = p.w; a.t0 = p.t0;
		for (int c = 0; c < 3; c++) { a.Ha0[c] = p.Ha[c]; a.dir[c] = p.dir[c]; a.slope[c] = p.slope[c]; }
	}
	if (o->ml[b]) {
		a.layered = 1;
		a.ihz2 = 0.f;
		a.lay = o->d_laytab + b;
		a.lay_grouped = o->lay_grouped[b];
		a.lay_field = o->lay_field[b];
		if (o->Msmap && (o->lay_set[b] & 8)) a.Ms_map = o->Msmap + o->off[b];
	}
	return a;
}

__global__ void k_time_advance(double* t, const double* dT) { *t += *dT; }

// Applied field of every mesh with a time-dependent one at its current step time, one thread
// per mesh, once per step (attempt), ahead of the stage kernels that read it.
__global__ void k_batch_fields(const StageArgs* __restrict__ tab, int nb)
{
	const int b = blockIdx.x * blockDim.x + threadIdx.x;
	if (b >= nb) return;
	const StageArgs& a = tab[b];
	if (!a.tnow || !a.Ha_dev) return;
	const F3 H = waveform_Ha(a);
	float* o = const_cast<float*>(a.Ha_dev);
	o[0] = H.x; o[1] = H.y; o[2] = H.z;
}

// RKDP54 step control after stage 0, one thread per mesh: the plain integrator's stage-0
// branch (primed check, SetAdaptiveTimeStep, restore on rejection) on the device.  The new
// dT goes into both stage tables (read in place by the stage kernels of this replay).
__global__ void k_rkdp_control(BatchCtl* ctl, unsigned* lte, StageArgs* t0, StageArgs* t1, int nb)
{
	const int b = blockIdx.x * blockDim.x + threadIdx.x;
	if (b >= nb) return;
	BatchCtl& c = ctl[b];
	if (!c.active) { c.skip = 1; return; }
	c.skip = 0;
	if (c.primed) {
		c.lte = (double)__uint_as_float(lte[b]);
		if (c.dT_min != c.dT_max) {
			if (c.lte > c.err_high_fail && c.dT > c.dT_min) {
				c.time -= c.dT; c.iteration--;
				double f = pow(c.err_high_fail * 0.8 / c.lte, 1.0 / (c.order + 1));
				if (f < 0.01) f = 0.01;
				c.dT *= f;
				if (c.dT < c.dT_min) c.dT = c.dT_min;
				c.rejections++;
				c.primed = 0;
				c.skip = 1;
			}
			else if (c.lte > 0) {
				double f = pow(c.err_high_fail * 0.8 / c.lte, 1.0 / (c.order + 1));
				if (f > c.dT_increase) f = c.dT_increase;
				if (f < 0.01) f = 0.01;
				c.dT *= f;
				if (c.dT < c.dT_min) c.dT = c.dT_min;
				if (c.dT > c.dT_max) c.dT = c.dT_max;
			}
		}
	}
	else c.primed = 1;
	lte[b] = 0u;
	t0[b].dT = (float)c.dT;
	t1[b].dT = (float)c.dT;
}

// After the control: an accepted mesh advances (sM1 = M, M1 = M + E0 dT/5 into the second
// array, which stage 1 reads); a rejected one restores M = sM1.  One block per flat chunk of
// kBS cells (3 kBS floats), mesh b owning chunks choff[b] .. choff[b + 1] - 1.
__global__ void __launch_bounds__(kBS) k_batch_rkdp_advance(const StageArgs* __restrict__ tab, const BatchCtl* __restrict__ ctl,
	const int* __restrict__ choff, int nb)
{
	const int ch = xcd_chunk(blockIdx.x, gridDim.x);
	if (ch >= choff[nb]) return;
	int lo = 0, hi = nb - 1;
	while (lo < hi) { const int mid = (lo + hi + 1) >> 1; if (choff[mid] <= ch) lo = mid; else hi = mid - 1; }
	const StageArgs& a = tab[lo];
	const BatchCtl& c = ctl[lo];
	if (!c.active) return;
	const size_t n3 = (size_t)a.nx * a.ny * a.nz * 3, q0 = (size_t)(ch - choff[lo]) * 3 * kBS + threadIdx.x;
	float* M = const_cast<float*>(a.Min);
	const float dT5 = (float)c.dT / 5;
#pragma unroll
	for (int r = 0; r < 3; r++) {
		const size_t q = q0 + (size_t)r * kBS;
		if (q >= n3) break;
		if (c.skip) { M[q] = a.sM1[q]; continue; }
		const float m = M[q];
		a.sM1[q] = m;
		a.Mout[q] = m + a.E[0][q] * dT5;
	}
}

// End of an attempt: an accepted step advances the mesh's time and iteration.
__global__ void k_rkdp_end(BatchCtl* ctl, int nb)
{
	const int b = blockIdx.x * blockDim.x + threadIdx.x;
	if (b >= nb) return;
	BatchCtl& c = ctl[b];
	if (!c.active || c.skip) return;
	c.time += c.dT;
	c.iteration++;
	if (c.iteration >= c.target) c.active = 0;
}

// Group the meshes by pass size.
void build_groups(bm_batch_s* o)
{
	const int nb = o->nb;
	for (int b = 0; b < nb; b++) {
		if (o->ml[b]) {
			const MlView v = mlayer_view(o->ml[b]);
			add_to(o->gx, o->nx[b], false, b);
			add_to(o->gy, v.Ny, false, b);
			add_to(o->gml, v.L, false, b);
			continue;
		}
		const bm_demag_s* d = o->dm[b];
		if (!d) continue;
		// a periodic x pass transforms rows of N.x = nx cells as nx/2-point complex FFTs
		const bool fx = d->pbc[0] > 0, fy = d->pbc[1] > 0, fz = d->pbc[2] > 0;
		add_to(o->gx, fx ? d->Nx / 2 : o->nx[b], fx, b);
		if (o->nz[b] > 1 && demag_yzy(d)) add_to(o->gyz[d->Nz == 8 ? 1 : 0], d->Ny, false, b);
		else if (o->nz[b] > 1) { add_to(o->gy, d->Ny, fy, b); add_to(o->gz, d->Nz, fz, b); }
		else add_to(o->gy2, d->Ny, fy, b);
	}
}

// Build every table on the host (x tables depend on the groups' tile factors) and upload
// them in one allocation, replacing earlier tables.
int build_tables(bm_batch_s* o)
{
	const int nb = o->nb;
	if (o->dev_tables) { (void)hipFree(o->dev_tables); o->dev_tables = nullptr; }
	for (int par = 0; par < 2; par++) o->fwd[par].assign(o->gx.size(), Tab<FwdArgs>());
	o->inv.assign(o->gx.size(), Tab<PassBatch>());
	o->ty.assign(o->gy.size(), Tab<PassBatch>());
	o->tz.assign(o->gz.size(), Tab<PassBatch>());
	o->ty2.assign(o->gy2.size(), Tab<PassBatch>());
	o->tml.assign(o->gml.size(), Tab<MlBatch>());
	for (int q = 0; q < 2; q++) {
		o->tyz[q].assign(o->gyz[q].size(), Tab<PassBatch>());
		for (size_t g = 0; g < o->gyz[q].size(); g++)
			for (int b : o->gyz[q][g].meshes) {
				bm_demag_s* d = o->dm[b];
				PassBatch pb{}; pb.Sx = d->Sx; pb.K = d->K; pb.nz = d->nz; pb.NC = d->NC; pb.NYH = d->NYH; pb.ntl = d->ntl; pb.ny = d->ny;
				o->tyz[q][g].h.push_back(pb);
			}
	}
	for (size_t g = 0; g < o->gx.size(); g++) {
		const int nx = o->gx[g].size;
		// the full-length (periodic) x passes run one row-tile height only
		const bool full = o->gx[g].full;
		const int tf = x_tile_rows(nx, true, full ? 1 : o->gx[g].sh[0].f), ti = x_tile_rows(nx, false, full ? 1 : o->gx[g].sh[1].f);
		for (int b : o->gx[g].meshes) {
			// a layer stack runs the x passes over its L layer planes
			float2* Sx; int ny, nz, ntl; double scale;
			if (o->ml[b]) { const MlView v = mlayer_view(o->ml[b]); Sx = v.Sx; ny = v.ny; nz = v.L; ntl = v.ntl; scale = 1.0 / ((double)v.Nx * v.Ny); }
			else { bm_demag_s* d = o->dm[b]; Sx = d->Sx; ny = d->ny; nz = d->nz; ntl = d->ntl; scale = 1.0 / ((double)d->Nx * d->Ny * d->Nz); }
			for (int par = 0; par < 2; par++) {
				FwdArgs f{};
				f.M = (par ? o->Mb : o->M) + 3 * o->off[b]; f.Sx = Sx;
				f.ny = ny; f.nz = nz; f.ntl = ntl; f.ntiles = (ny + tf - 1) / tf;
				o->fwd[par][g].h.push_back(f);
			}
			PassBatch q{};
			q.Sx = Sx; q.out = o->Hd + 3 * o->off[b]; q.ny = ny; q.nz = nz; q.ntl = ntl; q.ntiles = (ny + ti - 1) / ti;
			q.scale = (float)scale;
			o->inv[g].h.push_back(q);
		}
	}
	for (size_t g = 0; g < o->gy.size(); g++)
		for (int b : o->gy[g].meshes) {
			PassBatch q{};
			if (o->ml[b]) { const MlView v = mlayer_view(o->ml[b]); q.S = v.S; q.Sx = v.Sx; q.Sp = v.Sp; q.NC = v.NC; q.ntl = v.ntl; q.nlines = 3 * v.L * v.NC; q.ny = v.ny; }
			else { bm_demag_s* d = o->dm[b]; q.S = d->S; q.Sx = d->Sx; q.Sp = d->Sp; q.NC = d->NC; q.ntl = d->ntl; q.nlines = 3 * d->nz * d->NC; q.ny = d->ny; }
			o->ty[g].h.push_back(q);
		}
	for (size_t g = 0; g < o->gml.size(); g++)
		for (int b : o->gml[g].meshes) {
			const MlView v = mlayer_view(o->ml[b]);
			o->tml[g].h.push_back(MlBatch{ v.S, v.K, v.NC, v.Ny, v.Sp });
		}
	for (size_t g = 0; g < o->gz.size(); g++)
		for (int b : o->gz[g].meshes) {
			bm_demag_s* d = o->dm[b];
			PassBatch q{}; q.S = d->S; q.K = d->K; q.nz = d->nz; q.NC = d->NC; q.Ny = d->Ny; q.NYH = d->NYH; q.Sp = d->Sp;
			o->tz[g].h.push_back(q);
		}
	for (size_t g = 0; g < o->gy2.size(); g++)
		for (int b : o->gy2[g].meshes) {
			bm_demag_s* d = o->dm[b];
			PassBatch q{}; q.Sx = d->Sx; q.K = d->K; q.NC = d->NC; q.NYH = d->NYH; q.ntl = d->ntl; q.ny = d->ny;
			o->ty2[g].h.push_back(q);
		}
	for (int par = 0; par < 2; par++) {
		o->stage[par].h.clear();
		for (int b = 0; b < nb; b++) o->stage[par].h.push_back(stage_args(o, b, par));
	}
	o->choff.h.assign(1, 0);
	for (int b = 0; b < nb; b++) o->choff.h.push_back(o->choff.h.back() + (int)((o->ncell[b] + kBS - 1) / kBS));
	// one device allocation, 256-byte aligned pieces
	size_t bytes = 0;
	auto reserve = [&](size_t n) { size_t at = bytes; bytes += (n + 255) / 256 * 256; return at; };
	std::vector<std::pair<size_t, std::pair<const void*, size_t>>> pieces;
	auto place = [&](auto& t) {
		const size_t n = t.h.size() * sizeof(t.h[0]);
		const size_t at = reserve(n);
		pieces.push_back({ at, { (const void*)t.h.data(), n } });
		return at;
	};
	std::vector<size_t> at_fwd[2], at_inv, at_y, at_z, at_y2, at_ml;
	for (int par = 0; par < 2; par++) for (auto& t : o->fwd[par]) at_fwd[par].push_back(place(t));
	for (auto& t : o->inv) at_inv.push_back(place(t));
	for (auto& t : o->ty) at_y.push_back(place(t));
	for (auto& t : o->tz) at_z.push_back(place(t));
	for (auto& t : o->ty2) at_y2.push_back(place(t));
	for (auto& t : o->tml) at_ml.push_back(place(t));
	std::vector<size_t> at_yz[2];
	for (int q = 0; q < 2; q++) for (auto& t : o->tyz[q]) at_yz[q].push_back(place(t));
	const size_t at_s0 = place(o->stage[0]), at_s1 = place(o->stage[1]), at_ch = place(o->choff);
	if (hipMalloc(&o->dev_tables, std::max<size_t>(bytes, 256)) != hipSuccess) return fail(BM_ERR_NOMEM, "batch tables");
	o->dev_bytes = bytes;
	char* base = (char*)o->dev_tables;
	for (auto& pc : pieces) BM_HIP(hipMemcpy(base + pc.first, pc.second.first, pc.second.second, hipMemcpyHostToDevice));
	for (int par = 0; par < 2; par++) for (size_t g = 0; g < o->fwd[par].size(); g++) o->fwd[par][g].d = (FwdArgs*)(base + at_fwd[par][g]);
	for (size_t g = 0; g < o->inv.size(); g++) o->inv[g].d = (PassBatch*)(base + at_inv[g]);
	for (size_t g = 0; g < o->ty.size(); g++) o->ty[g].d = (PassBatch*)(base + at_y[g]);
	for (size_t g = 0; g < o->tz.size(); g++) o->tz[g].d = (PassBatch*)(base + at_z[g]);
	for (size_t g = 0; g < o->ty2.size(); g++) o->ty2[g].d = (PassBatch*)(base + at_y2[g]);
	for (size_t g = 0; g < o->tml.size(); g++) o->tml[g].d = (MlBatch*)(base + at_ml[g]);
	for (int q = 0; q < 2; q++) for (size_t g = 0; g < o->tyz[q].size(); g++) o->tyz[q][g].d = (PassBatch*)(base + at_yz[q][g]);
	o->stage[0].d = (StageArgs*)(base + at_s0);
	o->stage[1].d = (StageArgs*)(base + at_s1);
	o->choff.d = (int*)(base + at_ch);
	return BM_OK;
}

// rewrite dT in the stage tables (the only per-step-size data the kernels read) and the
// device copy the step-time advance reads
int upload_dt(bm_batch_s* o)
{
	BM_HIP(hipMemcpy(o->d_dT, &o->dT, sizeof(double), hipMemcpyHostToDevice));
	for (int par = 0; par < 2; par++) {
		for (StageArgs& a : o->stage[par].h) a.dT = (float)o->dT;
		BM_HIP(hipMemcpy(o->stage[par].d, o->stage[par].h.data(), o->stage[par].h.size() * sizeof(StageArgs), hipMemcpyHostToDevice));
	}
	return BM_OK;
}

// The n pass-size groups of one pass: launch(g, stream) for each.  With side streams the
// groups run as parallel branches (in a captured step: graph branches), joined on st before
// the next pass; each group's launch reads and writes only its own meshes' buffers.
template <typename F> int fan(bm_batch_s* o, size_t n, hipStream_t st, F&& launch)
{
	const int S = (int)std::min<size_t>(n, (size_t)o->nside + 1);
	int rc = BM_OK;
	if (S <= 1) {
		for (size_t g = 0; g < n && !rc; g++) rc = launch(g, st);
		return rc;
	}
	BM_HIP(hipEventRecord(o->ev_fork, st));
	for (int k = 1; k < S; k++) BM_HIP(hipStreamWaitEvent(o->side[k - 1], o->ev_fork, 0));
	for (size_t g = 0; g < n && !rc; g++) { const int k = (int)(g % S); rc = launch(g, k ? o->side[k - 1] : st); }
	for (int k = 1; k < S; k++) {
		BM_HIP(hipEventRecord(o->ev_join[k - 1], o->side[k - 1]));
		BM_HIP(hipStreamWaitEvent(st, o->ev_join[k - 1], 0));
	}
	return rc;
}

// One field evaluation + stage `step` for every mesh; input magnetisation M (parity 0) or Mb.
int eval_all(bm_batch_s* o, int step, int par, hipStream_t st)
{
	int rc = fan(o, o->gx.size(), st, [&](size_t g, hipStream_t s) {
		return batch_fwd_x(o->gx[g].size, o->gx[g].full, o->gx[g].sh[0], o->fwd[par][g].h.data(), o->fwd[par][g].d, (int)o->fwd[par][g].h.size(), o->ncu, s); });
	if (!rc) rc = fan(o, o->gy.size(), st, [&](size_t g, hipStream_t s) {
		return batch_y(o->gy[g].size, o->gy[g].full, -1, o->gy[g].sh[0], o->ty[g].h.data(), o->ty[g].d, (int)o->ty[g].h.size(), o->ncu, s); });
	if (!rc) rc = fan(o, o->gz.size(), st, [&](size_t g, hipStream_t s) {
		return batch_z(o->gz[g].size, o->gz[g].full, o->gz[g].sh[0], o->tz[g].h.data(), o->tz[g].d, (int)o->tz[g].h.size(), o->ncu, s); });
	if (!rc) rc = fan(o, o->gml.size(), st, [&](size_t g, hipStream_t s) {
		return batch_ml_mult(o->gml[g].size, o->gml[g].sh[0], o->tml[g].h.data(), o->tml[g].d, (int)o->tml[g].h.size(), o->ncu, s); });
	if (!rc) rc = fan(o, o->gy.size(), st, [&](size_t g, hipStream_t s) {
		return batch_y(o->gy[g].size, o->gy[g].full, 1, o->gy[g].sh[1], o->ty[g].h.data(), o->ty[g].d, (int)o->ty[g].h.size(), o->ncu, s); });
	if (!rc) rc = fan(o, o->gy2.size(), st, [&](size_t g, hipStream_t s) {
		return batch_y2d(o->gy2[g].size, o->gy2[g].full, o->gy2[g].sh[0], o->ty2[g].h.data(), o->ty2[g].d, (int)o->ty2[g].h.size(), o->ncu, s); });
	for (int q = 0; q < 2 && !rc; q++)
		rc = fan(o, o->gyz[q].size(), st, [&](size_t g, hipStream_t s) {
			return batch_yzy(o->gyz[q][g].size, q ? 8 : 4, o->gyz[q][g].sh[0], o->tyz[q][g].h.data(), o->tyz[q][g].d, (int)o->tyz[q][g].h.size(),
				o->ncu, s); });
	if (!rc) rc = fan(o, o->gx.size(), st, [&](size_t g, hipStream_t s) {
		return batch_inv_x_planar(o->gx[g].size, o->gx[g].full, o->gx[g].sh[1], o->inv[g].h.data(), o->inv[g].d, (int)o->inv[g].h.size(), o->ncu, s); });
	if (!rc) rc = batch_stage(o->mcode, step, o->stage[par].d, o->nb, o->choff.d, o->choff.h.back(), st);
	return rc;
}

// Autotuner: for every pass-size group and pass, time each candidate launch shape (five
// launches after a warm-up, HIP events) on the group's real tables and keep the fastest.
// Candidates: x passes tile factor {1, 2} x workgroups per CU {1, 2, 4}, and the half-height
// tile at {2, 4, 8} per CU where the row length has one; y passes {4, 8, 16};
// z pass {1, 2, 4}; 2D y pass {4, 8, 16}; multilayer product {4, 8, 16}.  Only the FFT passes
// run (their inputs are recomputed from M by every evaluation), so the state is untouched.
int autotune(bm_batch_s* o, hipStream_t st)
{
	hipEvent_t e0 = nullptr, e1 = nullptr;
	BM_HIP(hipEventCreate(&e0));
	BM_HIP(hipEventCreate(&e1));
	int rc = BM_OK;
	auto timed = [&](auto launch) -> double {
		if (launch() != BM_OK) return -1;
		(void)hipEventRecord(e0, st);
		for (int r = 0; r < 5; r++) if (launch() != BM_OK) return -1;
		(void)hipEventRecord(e1, st);
		if (hipEventSynchronize(e1) != hipSuccess) return -1;
		float ms = 0;
		(void)hipEventElapsedTime(&ms, e0, e1);
		return 1e3 * ms / 5;
	};
	auto pick = [&](Group& G, int slot, const std::vector<LaunchShape>& cands, bool retable, auto launch) {
		double bt = 1e30;
		LaunchShape best = cands[0];
		for (const LaunchShape& c : cands) {
			G.sh[slot] = c;
			if (retable && (rc = build_tables(o))) return;
			const double t = timed(launch);
			if (t >= 0 && t < bt) { bt = t; best = c; }
		}
		G.sh[slot] = best; G.us[slot] = bt;
	};
	std::vector<LaunchShape> cx, cxh, cy, cz;
	for (int f : { 1, 2 }) for (int pc : { 1, 2, 4 }) cx.push_back(LaunchShape{ f, pc });
	// with the half-height tile (twice the work items for small meshes) where the length has one
	cxh = cx;
	for (int pc : { 2, 4, 8 }) cxh.push_back(LaunchShape{ -2, pc });
	for (int pc : { 4, 8, 16 }) cy.push_back(LaunchShape{ 0, pc });
	for (int pc : { 1, 2, 4 }) cz.push_back(LaunchShape{ 0, pc });
	// the launch shapes crossed with every radix schedule of the pass length n
	std::vector<LaunchShape> cx1;
	for (int pc : { 1, 2, 4 }) cx1.push_back(LaunchShape{ 1, pc });
	// the launch shapes crossed with every radix schedule of the pass length n (periodic
	// groups: the default schedule only)
	auto radices = [&](const std::vector<LaunchShape>& base, int n, bool full) {
		std::vector<LaunchShape> r;
		const int nv = o->tune_radix && !full ? radix_variants(n) : 1;
		for (int v = 0; v < nv; v++) for (LaunchShape c : base) { c.radix = v; r.push_back(c); }
		return r;
	};
	for (size_t g = 0; g < o->gx.size() && !rc; g++) {
		Group& G = o->gx[g];
		const std::vector<LaunchShape>& cands = G.full ? cx1 : x_half_tile_ok(G.size) ? cxh : cx;
		pick(G, 0, cands, true, [&] { return batch_fwd_x(G.size, G.full, G.sh[0], o->fwd[0][g].h.data(), o->fwd[0][g].d, (int)o->fwd[0][g].h.size(), o->ncu, st); });
		if (!rc) pick(G, 1, cands, true, [&] { return batch_inv_x_planar(G.size, G.full, G.sh[1], o->inv[g].h.data(), o->inv[g].d, (int)o->inv[g].h.size(), o->ncu, st); });
	}
	for (size_t g = 0; g < o->gy.size() && !rc; g++) {
		Group& G = o->gy[g];
		for (int slot = 0; slot < 2; slot++)
			pick(G, slot, radices(cy, G.size, G.full), false, [&] { return batch_y(G.size, G.full, slot ? 1 : -1, G.sh[slot], o->ty[g].h.data(), o->ty[g].d, (int)o->ty[g].h.size(), o->ncu, st); });
	}
	for (size_t g = 0; g < o->gz.size() && !rc; g++) {
		Group& G = o->gz[g];
		pick(G, 0, radices(cz, G.size, G.full), false, [&] { return batch_z(G.size, G.full, G.sh[0], o->tz[g].h.data(), o->tz[g].d, (int)o->tz[g].h.size(), o->ncu, st); });
	}
	for (size_t g = 0; g < o->gy2.size() && !rc; g++) {
		Group& G = o->gy2[g];
		pick(G, 0, radices(cy, G.size, G.full), false, [&] { return batch_y2d(G.size, G.full, G.sh[0], o->ty2[g].h.data(), o->ty2[g].d, (int)o->ty2[g].h.size(), o->ncu, st); });
	}
	for (size_t g = 0; g < o->gml.size() && !rc; g++) {
		Group& G = o->gml[g];
		pick(G, 0, cy, false, [&] { return batch_ml_mult(G.size, G.sh[0], o->tml[g].h.data(), o->tml[g].d, (int)o->tml[g].h.size(), o->ncu, st); });
	}
	for (int q = 0; q < 2 && !rc; q++)
		for (size_t g = 0; g < o->gyz[q].size() && !rc; g++) {
			Group& G = o->gyz[q][g];
			pick(G, 0, cy, false, [&] { return batch_yzy(G.size, q ? 8 : 4, G.sh[0], o->tyz[q][g].h.data(), o->tyz[q][g].d, (int)o->tyz[q][g].h.size(), o->ncu, st); });
		}
	if (!rc) rc = build_tables(o);
	(void)hipEventDestroy(e0);
	(void)hipEventDestroy(e1);
	if (!rc) o->tuned = true;
	return rc;
}

int one_step(bm_batch_s* o, hipStream_t st)
{
	const int nbk = (o->nb + 63) / 64;
	// time-dependent applied fields at each mesh's step time, once per step or attempt (the
	// field is held at the step start time, so stages 1.. of an attempt see the same value)
	if (o->has_wave) hipLaunchKernelGGL(k_batch_fields, dim3(nbk), dim3(64), 0, st, (const StageArgs*)o->stage[0].d, o->nb);
	if (o->mcode == M_RKDP) {
		int rc = eval_all(o, 0, 0, st);
		if (rc) return rc;
		hipLaunchKernelGGL(k_rkdp_control, dim3(nbk), dim3(64), 0, st, o->d_ctl, o->d_lte, o->stage[0].d, o->stage[1].d, o->nb);
		const int nchunks = o->choff.h.back();
		hipLaunchKernelGGL(k_batch_rkdp_advance, dim3((unsigned)(((std::max(nchunks, 1) + 7) / 8) * 8)), dim3(kBS), 0, st,
			(const StageArgs*)o->stage[0].d, (const BatchCtl*)o->d_ctl, (const int*)o->choff.d, o->nb);
		for (int s = 1; s < 6; s++) if ((rc = eval_all(o, s, s & 1, st))) return rc;
		hipLaunchKernelGGL(k_rkdp_end, dim3(nbk), dim3(64), 0, st, o->d_ctl, o->nb);
		return BM_OK;
	}
	for (int s = 0; s < o->nstages; s++) {
		int rc = eval_all(o, s, s & 1, st);
		if (rc) return rc;
	}
	hipLaunchKernelGGL(k_time_advance, dim3(1), dim3(1), 0, st, o->d_time, o->d_dT);
	return BM_OK;
}

} // namespace

extern "C" int bm_batch_create(int method, int nmesh, const int* n3, const double* h3, const bm_fm_params* params,
	float* M, int flags, bm_batch_t* out)
{
	return bm_batch_create_layers(method, nmesh, n3, h3, params, nullptr, M, flags, out);
}

extern "C" int bm_batch_create_layers(int method, int nmesh, const int* n3, const double* h3, const bm_fm_params* params,
	const bm_layer_desc* layers, float* M, int flags, bm_batch_t* out)
{
	if (!out || nmesh < 1 || !n3 || !h3 || !params || !M) return fail(BM_ERR_ARG, "bm_batch_create: bad arguments");
	*out = nullptr;
	bm_batch_s* o = new bm_batch_s();
	o->nb = nmesh;
	o->use_graph = (flags & BM_BATCH_NO_GRAPH) ? 0 : 1;
	o->tune_radix = (flags & BM_BATCH_AUTOTUNE_RADIX) != 0;
	o->autotune = (flags & BM_BATCH_AUTOTUNE) != 0 || o->tune_radix;
	switch (method) {
	case BM_EVAL_RK4: o->mcode = M_RK4; o->nstages = 4; o->dT = 0.5e-12; break;
	case BM_EVAL_TEULER: o->mcode = M_TEULER; o->nstages = 2; o->dT = 0.5e-13; break;
	case BM_EVAL_RKDP54: o->mcode = M_RKDP; o->nstages = 6; o->dT = 0.5e-12; break;
	default: delete o; return fail(BM_ERR_ARG, "bm_batch_create: RKDP54, RK4 or TEuler");
	}
	o->M = M;
	o->ncu = device_cu_count();
	for (int b = 0; b < nmesh; b++) {
		const int nx = n3[3 * b], ny = n3[3 * b + 1], nz = n3[3 * b + 2];
		if (nx < 1 || ny < 1 || nz < 1) { delete o; return fail(BM_ERR_ARG, "bm_batch_create: bad mesh size"); }
		o->nx.push_back(nx); o->ny.push_back(ny); o->nz.push_back(nz);
		o->off.push_back(o->ntot);
		o->ncell.push_back((size_t)nx * ny * nz);
		o->ntot += (size_t)nx * ny * nz;
		o->max_cells = std::max(o->max_cells, (size_t)nx * ny * nz);
		for (int c = 0; c < 3; c++) o->h.push_back(h3[3 * b + c]);
		o->p.push_back(params[b]);
	}
	o->dm.assign(nmesh, nullptr);
	o->ml.assign(nmesh, nullptr);
	o->pbc.assign(nmesh, std::array<int, 3>{ 0, 0, 0 });
	o->lay.assign(nmesh, std::vector<float>(3 * kMaxStackLayers, 0.f));
	o->lay_set.assign(nmesh, 0);
	o->lay_grouped.assign(nmesh, 0);
	o->lay_field.assign(nmesh, 0);
	int rc = BM_OK;
	bool any_lay_ms = false;
	for (int b = 0; b < nmesh && rc == BM_OK; b++) {
		const bm_layer_desc* ld = layers ? &layers[b] : nullptr;
		if (ld && ld->nlayers > 0) {
			if (ld->nlayers != o->nz[b] || ld->nlayers > kMaxBatchStackLayers || !ld->hz || !ld->z0)
				rc = fail(BM_ERR_ARG, "bm_batch_create_layers: a stack needs nz == nlayers <= 16 and its hz, z0");
			else if (!(o->p[b].modules & BM_MOD_DEMAG))
				rc = fail(BM_ERR_ARG, "bm_batch_create_layers: a layer stack needs the demag module");
			else {
				rc = bm_mlayer_create(ld->nlayers, o->nx[b], o->ny[b], o->h[3 * b], o->h[3 * b + 1], ld->hz, ld->z0, 0, &o->ml[b]);
				if (rc == BM_OK && mlayer_mapped(o->ml[b])) rc = fail(BM_ERR_SIZE, "bm_batch_create_layers: a stack needs nx and 2 ny with FFT plans");
			}
			const double* src[3] = { ld->A, ld->alpha, ld->K1 };
			for (int q = 0; q < 3; q++)
				if (src[q]) { o->lay_set[b] |= 1 << q; for (int l = 0; l < ld->nlayers; l++) o->lay[b][q * kMaxStackLayers + l] = (float)src[q][l]; }
			if (ld->Ms) { o->lay_set[b] |= 8; any_lay_ms = true; }
		}
		else if (o->p[b].modules & BM_MOD_DEMAG) {
			rc = bm_demag_create(o->nx[b], o->ny[b], o->nz[b], o->h[3 * b], o->h[3 * b + 1], o->h[3 * b + 2], 0, &o->dm[b]);
			// batched x passes are grouped by row length and run the exact-length variant
			if (rc == BM_OK && o->dm[b]->Nx != 2 * o->nx[b])
				rc = fail(BM_ERR_SIZE, "bm_batch_create: batched meshes need nx in the 2^a 3^b plans 2..2048 (other nx run on single-mesh handles)");
		}
	}
	// per-layer tables of the stacks (A, alpha, K1 given per layer or the mesh's)
	if (rc == BM_OK && layers) {
		std::vector<LayerTab> tab(nmesh);
		std::memset(tab.data(), 0, tab.size() * sizeof(LayerTab));
		for (int b = 0; b < nmesh; b++)
			for (int l = 0; l < kMaxStackLayers; l++) {
				tab[b].A[l] = (o->lay_set[b] & 1) ? o->lay[b][l] : (float)o->p[b].A;
				tab[b].alpha[l] = (o->lay_set[b] & 2) ? o->lay[b][kMaxStackLayers + l] : (float)o->p[b].alpha;
				tab[b].K1[l] = (o->lay_set[b] & 4) ? o->lay[b][2 * kMaxStackLayers + l] : (float)o->p[b].K1;
			}
		// stacks of 3D meshes (as bm_ode_set_layer_meshes): planes of a mesh touch and share a thickness
		for (int b = 0; b < nmesh && rc == BM_OK; b++) {
			const bm_layer_desc& ld = layers[b];
			if (ld.nlayers <= 0) continue;
			if (ld.Ha && o->p[b].wave != 0) { rc = fail(BM_ERR_ARG, "bm_batch_create_layers: per-layer fields are constant"); break; }
			for (int l = 0; l < ld.nlayers; l++) {
				tab[b].grp[l] = ld.mesh_of_layer ? ld.mesh_of_layer[l] : l;
				tab[b].ihz2[l] = (float)(1.0 / (ld.hz[l] * ld.hz[l]));
				if (ld.Ha) for (int c = 0; c < 3; c++) tab[b].Ha[l][c] = (float)ld.Ha[3 * l + c];
				if (l > 0 && ld.mesh_of_layer) {
					const int g0 = ld.mesh_of_layer[l - 1], g1 = ld.mesh_of_layer[l];
					const double gap = ld.z0[l] - ld.z0[l - 1];
					if (g1 < g0 || (g1 == g0 && (std::fabs(ld.hz[l] - ld.hz[l - 1]) > 1e-6 * ld.hz[l] || std::fabs(gap - ld.hz[l - 1]) > 1e-6 * ld.hz[l])))
						rc = fail(BM_ERR_ARG, "bm_batch_create_layers: the planes of a mesh must be consecutive, touching and of one thickness");
				}
			}
			o->lay_grouped[b] = ld.mesh_of_layer ? 1 : 0;
			o->lay_field[b] = ld.Ha ? 1 : 0;
		}
		if (rc == BM_OK && (hipMalloc(&o->d_laytab, tab.size() * sizeof(LayerTab)) != hipSuccess
			|| hipMemcpy(o->d_laytab, tab.data(), tab.size() * sizeof(LayerTab), hipMemcpyHostToDevice) != hipSuccess))
			rc = fail(BM_ERR_NOMEM, "bm_batch_create_layers: layer tables");
	}
	if (rc == BM_OK && (hipMalloc(&o->d_time, sizeof(double)) != hipSuccess || hipMalloc(&o->d_dT, sizeof(double)) != hipSuccess
		|| hipMemset(o->d_time, 0, sizeof(double)) != hipSuccess || hipMemcpy(o->d_dT, &o->dT, sizeof(double), hipMemcpyHostToDevice) != hipSuccess))
		rc = fail(BM_ERR_NOMEM, "bm_batch_create: time scalars");
	for (int b = 0; b < nmesh; b++)
		if ((o->p[b].modules & BM_MOD_ZEEMAN) && (o->p[b].wave == 1 || o->p[b].wave == 2)) o->has_wave = true;
	if (rc == BM_OK && o->has_wave && (hipMalloc(&o->d_Ha, 3 * (size_t)nmesh * sizeof(float)) != hipSuccess
		|| hipMemset(o->d_Ha, 0, 3 * (size_t)nmesh * sizeof(float)) != hipSuccess))
		rc = fail(BM_ERR_NOMEM, "bm_batch_create: applied fields");
	if (rc == BM_OK && any_lay_ms) {
		std::vector<float> hm(o->ntot, 0.f);
		for (int b = 0; b < nmesh; b++) {
			if (!(o->lay_set[b] & 8)) continue;
			const size_t nl = (size_t)o->nx[b] * o->ny[b];
			for (size_t i = 0; i < o->ncell[b]; i++) hm[o->off[b] + i] = (float)layers[b].Ms[i / nl];
		}
		if (hipMalloc(&o->Msmap, o->ntot * sizeof(float)) != hipSuccess
			|| hipMemcpy(o->Msmap, hm.data(), o->ntot * sizeof(float), hipMemcpyHostToDevice) != hipSuccess)
			rc = fail(BM_ERR_NOMEM, "bm_batch_create: Ms map");
	}
	const size_t b3 = o->ntot * 3 * sizeof(float);
	if (rc == BM_OK) {
		bool ok = hipMalloc(&o->Mb, b3) == hipSuccess && hipMalloc(&o->sM1, b3) == hipSuccess && hipMalloc(&o->Hd, b3) == hipSuccess;
		const int nE = o->mcode == M_RK4 ? 3 : o->mcode == M_RKDP ? 6 : 0;
		for (int q = 0; q < nE && ok; q++) ok = hipMalloc(&o->E[q], b3) == hipSuccess && hipMemset(o->E[q], 0, b3) == hipSuccess;
		if (ok) ok = hipMemset(o->Hd, 0, b3) == hipSuccess && hipMemset(o->sM1, 0, b3) == hipSuccess;
		if (ok && o->mcode == M_RKDP) {
			// the reference's RKDP54 defaults (Boris/DiffEq_CommonBase_Control.cpp:10-140)
			ok = hipHostMalloc(&o->h_ctl, nmesh * sizeof(BatchCtl)) == hipSuccess
				&& hipMalloc(&o->d_ctl, nmesh * sizeof(BatchCtl)) == hipSuccess
				&& hipMalloc(&o->d_lte, nmesh * sizeof(unsigned)) == hipSuccess
				&& hipMemset(o->d_lte, 0, nmesh * sizeof(unsigned)) == hipSuccess;
			if (ok) {
				for (int b = 0; b < nmesh; b++) {
					BatchCtl& c = o->h_ctl[b];
					std::memset(&c, 0, sizeof(c));
					c.dT = o->dT; c.err_high_fail = 1e-4; c.dT_increase = 2; c.dT_min = 1e-15; c.dT_max = 3e-12; c.order = 5;
				}
				ok = hipMemcpy(o->d_ctl, o->h_ctl, nmesh * sizeof(BatchCtl), hipMemcpyHostToDevice) == hipSuccess;
			}
		}
		if (!ok) rc = fail(BM_ERR_NOMEM, "bm_batch_create: device arrays");
	}
	if (rc == BM_OK) { build_groups(o); rc = build_tables(o); }
	if (rc == BM_OK && o->use_graph) {
		if (hipStreamCreate(&o->gs) != hipSuccess || hipEventCreateWithFlags(&o->ev_in, hipEventDisableTiming) != hipSuccess
			|| hipEventCreateWithFlags(&o->ev_out, hipEventDisableTiming) != hipSuccess) rc = fail(BM_ERR_HIP, "bm_batch_create: stream");
	}
	if (rc == BM_OK) {
		// concurrent branches per pass: 4 (64 uneven films: RK4 step 3.77 -> 3.12 ms against
		// one branch, 8 no better); BM_BATCH_STREAMS overrides (1 = the groups one after another)
		const char* e = getenv("BM_BATCH_STREAMS");
		const int nbr = std::max(1, std::min(e ? atoi(e) : 4, bm_batch_s::kMaxSide + 1));
		if (nbr > 1 && hipEventCreateWithFlags(&o->ev_fork, hipEventDisableTiming) != hipSuccess) rc = fail(BM_ERR_HIP, "bm_batch_create: events");
		for (int k = 0; k + 1 < nbr && rc == BM_OK; k++) {
			// non-blocking: no implicit ordering against the legacy default stream (the fork and
			// join events order the branches)
			if (hipStreamCreateWithFlags(&o->side[k], hipStreamNonBlocking) != hipSuccess
				|| hipEventCreateWithFlags(&o->ev_join[k], hipEventDisableTiming) != hipSuccess)
				rc = fail(BM_ERR_HIP, "bm_batch_create: side streams");
			o->nside = k + 1;
		}
	}
	if (rc) { free_all(o); delete o; return rc; }
	*out = o;
	return BM_OK;
}

extern "C" int bm_batch_destroy(bm_batch_t o)
{
	if (!o) return BM_OK;
	(void)hipDeviceSynchronize();
	free_all(o);
	delete o;
	return BM_OK;
}

// host copy of the control blocks -> device, and their dT into the stage tables
static int upload_ctl(bm_batch_s* o)
{
	BM_HIP(hipMemcpy(o->d_ctl, o->h_ctl, o->nb * sizeof(BatchCtl), hipMemcpyHostToDevice));
	for (int par = 0; par < 2; par++) {
		for (int b = 0; b < o->nb; b++) o->stage[par].h[b].dT = (float)o->h_ctl[b].dT;
		BM_HIP(hipMemcpy(o->stage[par].d, o->stage[par].h.data(), o->stage[par].h.size() * sizeof(StageArgs), hipMemcpyHostToDevice));
	}
	return BM_OK;
}

static int fetch_ctl(bm_batch_s* o, hipStream_t st)
{
	BM_HIP(hipMemcpyAsync(o->h_ctl, o->d_ctl, o->nb * sizeof(BatchCtl), hipMemcpyDeviceToHost, st));
	BM_HIP(hipStreamSynchronize(st));
	return BM_OK;
}

extern "C" int bm_batch_set_dt(bm_batch_t o, double dT)
{
	if (!o || !(dT > 0)) return fail(BM_ERR_ARG, "bm_batch_set_dt: bad arguments");
	o->dT = dT;
	if (o->mcode == M_RKDP) {
		BM_HIP(hipDeviceSynchronize());
		for (int b = 0; b < o->nb; b++) o->h_ctl[b].dT = dT;
		return upload_ctl(o);
	}
	return upload_dt(o);
}

extern "C" int bm_batch_set_astep(bm_batch_t o, double err_fail, double dT_increase, double dT_min, double dT_max)
{
	if (!o) return fail(BM_ERR_ARG, "null handle");
	if (o->mcode != M_RKDP) return fail(BM_ERR_STATE, "bm_batch_set_astep: adaptive control is for RKDP54 batches");
	BM_HIP(hipDeviceSynchronize());
	for (int b = 0; b < o->nb; b++) {
		BatchCtl& c = o->h_ctl[b];
		if (err_fail > 0) c.err_high_fail = err_fail;
		if (dT_increase > 0) c.dT_increase = dT_increase;
		if (dT_min > 0) c.dT_min = dT_min;
		if (dT_max > 0) c.dT_max = dT_max;
	}
	return upload_ctl(o);
}

extern "C" int bm_batch_get_mesh_state(bm_batch_t o, int mesh, double* time, double* dT, long long* iteration, long long* rejections)
{
	if (!o || mesh < 0 || mesh >= o->nb) return fail(BM_ERR_ARG, "bm_batch_get_mesh_state: bad mesh");
	if (o->mcode != M_RKDP) {
		if (time) *time = o->time;
		if (dT) *dT = o->dT;
		if (iteration) *iteration = o->iteration;
		if (rejections) *rejections = 0;
		return BM_OK;
	}
	const BatchCtl& c = o->h_ctl[mesh];
	if (time) *time = c.time;
	if (dT) *dT = c.dT;
	if (iteration) *iteration = c.iteration;
	if (rejections) *rejections = c.rejections;
	return BM_OK;
}

// k replays of one step (one attempt per mesh for RKDP54): captured graph or direct launches
static int replay(bm_batch_s* o, long long k, hipStream_t st)
{
	int rc;
	if (!o->use_graph) {
		for (long long q = 0; q < k; q++) if ((rc = one_step(o, st))) return rc;
		BM_HIP(hipGetLastError());
		return BM_OK;
	}
	if (!o->exec) {
		// capture one step on the internal stream
		BM_HIP(hipStreamBeginCapture(o->gs, hipStreamCaptureModeThreadLocal));
		rc = one_step(o, o->gs);
		hipGraph_t g = nullptr;
		hipError_t e = hipStreamEndCapture(o->gs, &g);
		if (rc) { if (g) (void)hipGraphDestroy(g); return rc; }
		if (e != hipSuccess) return hip_fail(e, "bm_batch_run: capture");
		o->graph = g;
		BM_HIP(hipGraphInstantiate(&o->exec, o->graph, nullptr, nullptr, 0));
	}
	BM_HIP(hipEventRecord(o->ev_in, st));
	BM_HIP(hipStreamWaitEvent(o->gs, o->ev_in, 0));
	for (long long q = 0; q < k; q++) BM_HIP(hipGraphLaunch(o->exec, o->gs));
	BM_HIP(hipEventRecord(o->ev_out, o->gs));
	BM_HIP(hipStreamWaitEvent(st, o->ev_out, 0));
	return BM_OK;
}

// RKDP54: every mesh makes nsteps more accepted steps.  Attempts that a mesh rejects cost
// it a replay, so after the first nsteps replays the control blocks are read back and the
// replays repeat for the largest remaining count until every mesh is done.
static int run_adaptive(bm_batch_s* o, long long nsteps, hipStream_t st)
{
	if (nsteps == 0) return BM_OK;
	for (int b = 0; b < o->nb; b++) { o->h_ctl[b].target = o->h_ctl[b].iteration + nsteps; o->h_ctl[b].active = 1; }
	BM_HIP(hipMemcpyAsync(o->d_ctl, o->h_ctl, o->nb * sizeof(BatchCtl), hipMemcpyHostToDevice, st));
	long long remaining = nsteps, replays = 0;
	const long long cap = 64 * nsteps + 4096;     // each rejection shrinks dT by >= 100x towards dT_min
	int rc;
	while (remaining > 0) {
		if (replays > cap) return fail(BM_ERR_STATE, "bm_batch_run: step control did not converge");
		if ((rc = replay(o, remaining, st))) return rc;
		replays += remaining;
		if ((rc = fetch_ctl(o, st))) return rc;
		remaining = 0;
		for (int b = 0; b < o->nb; b++) remaining = std::max(remaining, o->h_ctl[b].target - o->h_ctl[b].iteration);
	}
	o->iteration += nsteps;
	o->time = o->h_ctl[0].time;
	o->dT = o->h_ctl[0].dT;
	return BM_OK;
}

extern "C" int bm_batch_run(bm_batch_t o, long long nsteps, void* stream)
{
	if (!o || nsteps < 0) return fail(BM_ERR_ARG, "bm_batch_run: bad arguments");
	hipStream_t st = (hipStream_t)stream;
	int rc;
	o->started = true;
	if (o->autotune && !o->tuned && nsteps > 0) {
		if (o->exec) { (void)hipGraphExecDestroy(o->exec); o->exec = nullptr; }
		if (o->graph) { (void)hipGraphDestroy(o->graph); o->graph = nullptr; }
		if ((rc = autotune(o, st))) return rc;
	}
	if (o->mcode == M_RKDP) return run_adaptive(o, nsteps, st);
	if ((rc = replay(o, nsteps, st))) return rc;
	for (long long k = 0; k < nsteps; k++) { o->time += o->dT; o->iteration++; }
	return BM_OK;
}

extern "C" int bm_batch_get_state(bm_batch_t o, double* time, long long* iteration)
{
	if (!o) return fail(BM_ERR_ARG, "null handle");
	if (time) *time = o->time;
	if (iteration) *iteration = o->iteration;
	return BM_OK;
}

extern "C" int bm_batch_mesh_offset(bm_batch_t o, int mesh, long long* cell_offset, long long* ncells)
{
	if (!o || mesh < 0 || mesh >= o->nb) return fail(BM_ERR_ARG, "bm_batch_mesh_offset: bad mesh");
	if (cell_offset) *cell_offset = (long long)o->off[mesh];
	if (ncells) *ncells = (long long)o->ncell[mesh];
	return BM_OK;
}

// The pass-size group and slot of a pass name ("x_fwd", "x_inv", "y_fwd", "y_inv", "z",
// "y2d", "ml_mult"; "_pbc" appended for a group along a periodic axis) and size, or nullptr.
static Group* find_group(bm_batch_s* o, std::string pass, int size, int& slot)
{
	std::vector<Group>* gs = nullptr;
	slot = 0;
	const std::string sfx = "_pbc";
	const bool full = pass.size() > sfx.size() && pass.compare(pass.size() - sfx.size(), sfx.size(), sfx) == 0;
	if (full) pass.resize(pass.size() - sfx.size());
	if (pass == "x_fwd" || pass == "x_inv") { gs = &o->gx; slot = pass == "x_inv"; }
	else if (pass == "y_fwd" || pass == "y_inv") { gs = &o->gy; slot = pass == "y_inv"; }
	else if (pass == "z") gs = &o->gz;
	else if (pass == "y2d") gs = &o->gy2;
	else if (pass == "ml_mult") gs = &o->gml;
	else if (pass == "yzy4") gs = &o->gyz[0];
	else if (pass == "yzy8") gs = &o->gyz[1];
	if (!gs) return nullptr;
	for (Group& G : *gs) if (G.size == size && G.full == full) return &G;
	return nullptr;
}

// One line per tuned pass:
// "<pass> <size> f=<factor> per_cu=<n> radix=<v> <schedule> <microseconds> us".
extern "C" int bm_batch_autotune_report(bm_batch_t o, char* out, int cap)
{
	if (!o || !out || cap < 1) return fail(BM_ERR_ARG, "bm_batch_autotune_report: bad arguments");
	std::string r;
	char line[192];
	auto add = [&](const char* pass, const Group& G, int slot) {
		const LaunchShape& c = G.sh[slot];
		const bool fft = (pass[0] == 'y' || pass[0] == 'z') && std::strncmp(pass, "yzy", 3) != 0;
		snprintf(line, sizeof(line), "%s%s %d f=%d per_cu=%d radix=%d %s %.1f us\n", pass, G.full ? "_pbc" : "", G.size, c.f, c.per_cu, c.radix,
			fft ? radix_schedule(G.size, c.radix) : "-", G.us[slot]);
		r += line;
	};
	if (o->tuned) {
		for (const Group& G : o->gx) { add("x_fwd", G, 0); add("x_inv", G, 1); }
		for (const Group& G : o->gy) { add("y_fwd", G, 0); add("y_inv", G, 1); }
		for (const Group& G : o->gz) add("z", G, 0);
		for (const Group& G : o->gy2) add("y2d", G, 0);
		for (const Group& G : o->gml) add("ml_mult", G, 0);
		for (const Group& G : o->gyz[0]) add("yzy4", G, 0);
		for (const Group& G : o->gyz[1]) add("yzy8", G, 0);
	}
	snprintf(out, (size_t)cap, "%s", r.c_str());
	return BM_OK;
}

extern "C" int bm_batch_set_schedule(bm_batch_t o, const char* pass, int size, int f, int per_cu, int radix)
{
	if (!o || !pass) return fail(BM_ERR_ARG, "bm_batch_set_schedule: bad arguments");
	int slot = 0;
	Group* G = find_group(o, pass, size, slot);
	if (!G) return fail(BM_ERR_ARG, "bm_batch_set_schedule: no such pass and size in this batch");
	// the fused thin-mesh pass has one schedule per length (radix 0 only)
	const bool x = pass[0] == 'x', fft = (pass[0] == 'y' || pass[0] == 'z') && std::strncmp(pass, "yzy", 3) != 0;
	const bool f_ok = f == 0 || f == 1 || (!G->full && (f == 2 || (f == -2 && x_half_tile_ok(size))));
	if ((x && !f_ok) || (!x && f != 0) || per_cu < 0 || per_cu > 64 || radix < 0
		|| (fft && !G->full ? radix >= radix_variants(size) : radix != 0))
		return fail(BM_ERR_ARG, "bm_batch_set_schedule: shape or radix schedule out of range for this pass");
	G->sh[slot] = LaunchShape{ f, per_cu, radix };
	G->us[slot] = 0;
	if (o->exec) { (void)hipGraphExecDestroy(o->exec); o->exec = nullptr; }
	if (o->graph) { (void)hipGraphDestroy(o->graph); o->graph = nullptr; }
	int rc;
	if (x && (rc = build_tables(o))) return rc;
	o->autotune = false;
	o->tuned = true;
	return BM_OK;
}

// Periodic boundary conditions for one mesh of the batch (the contract of bm_ode_set_pbc):
// its demag handle is rebuilt with the images (bm_demag_create_pbc: along a periodic axis
// N = n) and exchange wraps around.  Before the first bm_batch_run; the pass groups are
// rebuilt, so schedules pinned earlier return to their defaults.
extern "C" int bm_batch_set_pbc(bm_batch_t o, int mesh, int pbc_x, int pbc_y, int pbc_z)
{
	if (!o || mesh < 0 || mesh >= o->nb || pbc_x < 0 || pbc_y < 0 || pbc_z < 0) return fail(BM_ERR_ARG, "bm_batch_set_pbc: bad arguments");
	if (o->started) return fail(BM_ERR_STATE, "bm_batch_set_pbc: set periodic images before the first bm_batch_run");
	if (o->ml[mesh]) return fail(BM_ERR_STATE, "bm_batch_set_pbc: layer stacks have no periodic variant");
	const int b = mesh;
	if (o->p[b].modules & BM_MOD_DEMAG) {
		bm_demag_t d = nullptr;
		const double* h = &o->h[3 * b];
		int rc = (pbc_x || pbc_y || pbc_z) ? bm_demag_create_pbc(o->nx[b], o->ny[b], o->nz[b], h[0], h[1], h[2], pbc_x, pbc_y, pbc_z, 0, &d)
		                                   : bm_demag_create(o->nx[b], o->ny[b], o->nz[b], h[0], h[1], h[2], 0, &d);
		if (rc) return rc;
		if (!pbc_x && d->Nx != 2 * o->nx[b]) {
			bm_demag_destroy(d);
			return fail(BM_ERR_SIZE, "bm_batch_set_pbc: batched meshes need nx in the 2^a 3^b plans 2..2048");
		}
		bm_demag_destroy(o->dm[b]);
		o->dm[b] = d;
	}
	o->pbc[b] = { pbc_x, pbc_y, pbc_z };
	for (std::vector<Group>* gs : { &o->gx, &o->gy, &o->gz, &o->gy2, &o->gml, &o->gyz[0], &o->gyz[1] }) gs->clear();
	build_groups(o);
	if (o->tuned && !o->autotune) o->tuned = false;
	return build_tables(o);
}
