// Internal declarations shared by the library's translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <string>

#include "boris_mi355x.h"

namespace bm {

void set_error(const std::string& msg);
int fail(int code, const std::string& msg);
int hip_fail(hipError_t e, const char* what);

#define BM_HIP(call) do { hipError_t e_ = (call); if (e_ != hipSuccess) return ::bm::hip_fail(e_, #call); } while (0)

int device_cu_count();

// Store of a complex value with the non-temporal hint (streamed to memory rather than kept in
// the caches): for the y and z passes' outputs, which the same pass never reads again and
// which at the sizes that matter (hundreds of MB) outgrow the caches anyway.  Measured at
// 512x512x32: y inverse 131 -> 122 us, z 211 -> 208 us, and the row kernel that follows
// 254 -> 248 us; the same hint on 8-byte stores at a 24-byte stride (the fused advance's
// AoS writes) doubles that pass.
__device__ __forceinline__ void st_nt(float2* p, float x, float y)
{
	typedef float v2f __attribute__((ext_vector_type(2)));
	v2f v; v.x = x; v.y = y;
	__builtin_nontemporal_store(v, reinterpret_cast<v2f*>(p));
}

// Launch timing: when enabled (bm_prof_begin), every kernel launch of the hot path is
// bracketed by HIP events on its own stream and tagged with its kind.
enum ProfKind { PK_FWD_X = 0, PK_Y_FWD, PK_Z, PK_Y_INV, PK_INV_X, PK_Y2D, PK_STAGE, PK_ADVANCE, PK_ROW, PK_FWD_ADV, PK_ML_MULT, PK_COUNT };
void prof_mark_begin(int kind, hipStream_t st);
void prof_mark_end(int kind, hipStream_t st);
// self-demag coefficient (-N diagonal) of one cell of size hx x hy x hz (bm_tensor.cpp)
void self_demag_coeff(double hx, double hy, double hz, double out[3]);
void self_demag_coeff_pbc(double hx, double hy, double hz, int nx, int ny, int nz, int px, int py, int pz, double out[3]);

} // namespace bm

struct bm_demag_s {
	int nx, ny, nz;
	int Nx, Ny, Nz, NC, NYH, NZH;
	int pbc[3] = { 0, 0, 0 }; // periodic images per axis (0: open boundary, N = 2n or more)
	bool full[3] = { false, false, false };  // periodic axis: N = n, no zero-padded half
	int Sp;                   // line pitch of the full spectra S (complex elements, >= Ny)
	int ntl;                  // row tiles per plane of the x-side spectra Sx (bm_xpass.h)
	double h[3];
	bool is2d;
	float2* Sx = nullptr;     // x-side spectra, row-tiled: 3 * nz * ntl * NC * kLT complex
	float2* S = nullptr;      // full spectra S[c][z][kx][ky < Ny] (3D only): 3 * nz * NC * Sp complex
	float* K = nullptr;       // kernels (see bm_demag_copy_kernels)
	long long kcount = 0;
	double* ebuf = nullptr;   // energy partials (one per P5 work item)
	float* avg = nullptr;     // averaged-input convolution: (M1 + M2) / 2 and the field (AoS,
	float* hbuf = nullptr;    // allocated on first use)
	int ecap = 0;
	long long bytes = 0;
	int ncu = 256;
	bool row_wave = false;    // integrator row kernel at nx = 512: the wave-independent variant
	                          // (BM_ROW_WAVE=1 at creation; measured slower, kept for tests)
	int yzy = 1;              // fused y, z, y pass for thin meshes (N.z 4 or 8): 1 for N.z = 4 from
	                          // 128 x 128 cells per plane, 2 always, 0 never (BM_YZY at creation)
};

namespace bm {
// device address of the allocation-free averaged convolution's static scratch on the
// current device (resolved with hipGetSymbolAddress once per device)
int static_scratch(float** out);
// H_demag(M) into planar output H3n (3 planes of n floats)
int demag_planar(bm_demag_s* d, const float* M, float* H3n, hipStream_t st);
// the convolution's passes, for the integrator's fused evaluation
int demag_fwd_x(bm_demag_s* d, const float* M, hipStream_t st);
int demag_fwd_x_advance(bm_demag_s* d, const float* M, const float* E0, float* sM1, float* Mout, float dT5, hipStream_t st);
int demag_mid(bm_demag_s* d, hipStream_t st);
// single passes over "planes" of ny rows (a z plane or a layer) for other pipelines; the y
// passes go between the row-tiled Sx (ny rows, ntl tiles per plane) and the full spectra S
// (Ny rows, line pitch Sp), nlines = planes * NC lines.  nx: the x-pass size N.x / 2; full: the
// axis is periodic (x: rows of N.x = 2 nx cells; y: N.y data rows, no zero padding)
int pass_fwd_x(int nx, const float* M, float2* Sx, int ny, int nplanes, int ntl, int ncu, hipStream_t st, bool full = false);
int pass_y(int Ny, int dir, float2* S, float2* Sx, int nlines, int Sp, int NC, int ntl, int ncu, hipStream_t st, bool full = false);
int pass_inv_x(int nx, int mode, const float2* Sx, float* out, int ny, int nplanes, int ntl, float scale, int ncu, hipStream_t st,
	bool full = false);
// the x passes of plane `plane` of a spectra array of nzl planes per component (one layer
// of a multilayered convolution): M / out hold that plane only (nx * ny cells, AoS); the
// inverse sets (mode 1) or adds (mode 2) and, with energy set, writes -mu0/(2 nx ny) sum
// M.H of the plane into *energy using epart (x_plane_parts(nx, ny) doubles)
int pass_fwd_x_plane(int nx, const float* M, float2* Sx, int ny, int plane, int nzl, int ntl, int ncu, hipStream_t st, bool full = false);
int pass_inv_x_plane(int nx, int mode, const float2* Sx, float* out, int ny, int plane, int nzl, int ntl, float scale, int ncu,
	hipStream_t st, const float* M, double* epart, float* energy, bool full = false);
int x_plane_parts(int nx, int ny);
// FP64 real-input line FFTs of the kernel setup (see k_line_fft)
int setup_line_fft(int N, const double* src, double* dst, int nlines, int na, long long sa, long long sb, long long se,
	long long da, long long db, long long de, int keep, int part, double sign, hipStream_t st);
int setup_line_fft(int N, const double* src, float* dst, int nlines, int na, long long sa, long long sb, long long se,
	long long da, long long db, long long de, int keep, int part, double sign, hipStream_t st);
// are the x pass (nx), y pass (Ny) and setup x FFT (2 nx) plans available?
bool plan_2d_ok(int nx, int Ny);

// Per-mesh arguments of the x R2C pass (also its single-mesh launch arguments); with
// E0 set it performs the RKDP54 advance on the way (sM1 = M, Mout = M + E0 dT/5).
struct FwdArgs {
	const float* M; float2* Sx;
	int ny, nz, ntl, ntiles;
	const float* E0; float* sM1; float* Mout; float dT5;
	int nx;                 // cells per row (the general-length variant reads it)
	// one layer of a multilayered spectra array: planes z of M go to plane zoff + z of Sx,
	// whose layout has nzl planes per component (0, 0: the planes of M, nz per component)
	int zoff, nzl;
};
// Per-mesh arguments of the other passes in a batched launch (the fields each pass uses;
// the y passes take ny as the data rows of a line).
struct PassBatch {
	float2* S; float2* Sx; const float* K; float* out;
	int ny, nz, Sp, ntl, ntiles, nlines, NC, Ny, NYH;
	float scale;            // x C2R: 1 / (Nx Ny Nz) of the mesh
};
// multilayered demag (bm_mlayer.hip): H_demag of all layers into planar H3n (3 planes of
// L*nx*ny); the handle's layer count, in-plane size and layer thicknesses (hz: L values)
int mlayer_planar(bm_mlayer_s* m, const float* M, float* H3n, hipStream_t st);
// the transposed operator N^T (the adjoint of the multilayered convolution), planar output
int mlayer_planar_transposed(bm_mlayer_s* m, const float* M, float* H3n, hipStream_t st);
void mlayer_dims(const bm_mlayer_s* m, int* L, int* nx, int* ny, double* hz);
// the handle's layers are meshes of their own on a common grid (bm_mlayer_create_meshes, or a
// size without a plan): only the convolution calls take it
bool mlayer_mapped(const bm_mlayer_s* m);
// the smallest x-pass length (2^a 3^b, 2..2048) at or above v, 0 if none
int plan_length_at_least(int v);
// a 2D grid of nx x ny cells has plans with x (fx) and y (fy) periodic or open
bool plan_2d_pbc_ok(int nx, int ny, bool fx, bool fy);
// periodic images along x and y of a multilayer handle (0, 0: open)
void mlayer_pbc(const bm_mlayer_s* m, int out[2]);
// Pointers and sizes of a multilayered demag handle (batched launches).
struct MlView {
	float2* Sx; float2* S; const float* K;
	int L, nx, ny, Nx, Ny, NC, NYH, Sp, ntl;
};
MlView mlayer_view(bm_mlayer_s* m);
// cell size (m) of layer l of a multilayer handle
void mlayer_cell(const bm_mlayer_s* m, int l, double h[3]);
double mlayer_z0(const bm_mlayer_s* m, int l);
// Per-stack arguments of the batched multilayer kernel product.
struct MlBatch { float2* S; const float* K; int NC, Ny, Sp; };
// Launch shape of a batched pass (0 = the pass's default): f the x-pass row-tile factor (1 or
// 2 times the base tile, -2 half of it), per_cu the grid's workgroups per CU, radix the radix
// schedule (plan variant of the pass length).  Picked per pass-size group by the batched
// integrator's autotuner.
struct LaunchShape { int f = 0, per_cu = 0, radix = 0; };
// number of radix schedules of length n, and schedule v as text ("16x16x4")
int radix_variants(int n);
const char* radix_schedule(int n, int v);
int batch_ml_mult(int L, LaunchShape sh, const MlBatch* h_tab, const MlBatch* d_tab, int nb, int ncu, hipStream_t st);
// rows per work item of the x passes for nx (0 if there is no plan); f: 0 the default, 1 the
// base tile, 2 twice it, -2 half of it
int x_tile_rows(int nx, bool fwd, int f = 0);
// the x passes of length nx have a half-height tile (f = -2)
bool x_half_tile_ok(int nx);
// Batched passes over nb meshes sharing the pass size (tables: host copy for the grid size,
// device copy for the kernels).
int batch_fwd_x(int nx, bool full, LaunchShape sh, const FwdArgs* h_tab, const FwdArgs* d_tab, int nb, int ncu, hipStream_t st);
int batch_inv_x_planar(int nx, bool full, LaunchShape sh, const PassBatch* h_tab, const PassBatch* d_tab, int nb, int ncu, hipStream_t st);
int batch_y(int Ny, bool full, int dir, LaunchShape sh, const PassBatch* h_tab, const PassBatch* d_tab, int nb, int ncu, hipStream_t st);
int batch_z(int Nz, bool full, LaunchShape sh, const PassBatch* h_tab, const PassBatch* d_tab, int nb, int ncu, hipStream_t st);
int batch_y2d(int Ny, bool full, LaunchShape sh, const PassBatch* h_tab, const PassBatch* d_tab, int nb, int ncu, hipStream_t st);
// thin meshes (N.z 4 or 8, open y and z, a power-of-two N.y with a fitting schedule): the y, z,
// y passes as one kernel that never writes the full spectra; demag_yzy: the handle takes it
bool demag_yzy(const bm_demag_s* d);
int batch_yzy(int Ny, int Nz, LaunchShape sh, const PassBatch* h_tab, const PassBatch* d_tab, int nb, int ncu, hipStream_t st);
}
