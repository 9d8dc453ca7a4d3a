/*
 * boris_mi355x.h — C-ABI of the MI355X demag-convolution + LLG-stage hot path.
 *
 * Plain pointers and sizes only.  All device pointers are HIP device memory owned by the
 * caller unless stated otherwise; `stream` is a hipStream_t passed as void* (NULL = the
 * default stream).  Every call returns an int status (BM_OK = 0); bm_last_error() gives a
 * message for the calling thread's last failure.  No exceptions cross this boundary.
 *
 * Data layout (the reference's cuVEC<cuReal3> layout, BorisCUDALib/cuVEC.h:58-85):
 * a vector field over an nx*ny*nz mesh is float[3*n], AoS (x,y,z per cell), cell index
 * i + j*nx + k*nx*ny.  M is in A/m with |M| = Ms; fields are in A/m.
 */
#ifndef BORIS_MI355X_H
#define BORIS_MI355X_H

#ifdef __cplusplus
extern "C" {
#endif

enum {
	BM_OK = 0,
	BM_ERR_ARG = 1,      /* invalid argument */
	BM_ERR_SIZE = 2,     /* mesh size has no FFT plan in this library */
	BM_ERR_NOMEM = 3,    /* device allocation failed (BERROR_OUTOFGPUMEMORY_CRIT) */
	BM_ERR_HIP = 4,      /* HIP runtime error */
	BM_ERR_STATE = 5     /* call not valid in the object's current state */
};

const char* bm_last_error(void);
/* 1 if a device is visible and the kernels in this library can run on it. */
int bm_device_ok(void);

/* ------------------------------------------------------------------------------------
 * Demag convolution of one ferromagnetic mesh.
 *
 * Replaces: DemagCUDA::Initialize + ConvolutionDataCUDA::SetConvolutionDimensions +
 * DemagKernelCUDA::Calculate_Demag_Kernels (Boris/DemagCUDA.cpp:31,
 * Boris/ConvolutionDataCUDA.cpp:24, Boris/DemagKernelCUDA.cpp:58/248/358/603):
 * computes the FP64 Newell tensor and its FFT on the GPU once, stores the FP32 kernels
 * and allocates the spectral scratch.  Sizes: any nx <= 4096, ny <= 2048, nz <= 128; N.x and
 * N.y the smallest planned lengths >= 2nx, 2ny (the 2^a 3^b plans, N.y up to 4096: the same
 * zero-padded convolution as the reference's N = 2n), N.z = 1 (nz = 1) or the next power of
 * two >= 2nz, as the reference's GPU path does; larger meshes return BM_ERR_SIZE.
 */
typedef struct bm_demag_s* bm_demag_t;

int bm_demag_create(int nx, int ny, int nz, double hx, double hy, double hz, int flags, bm_demag_t* out);
/*
 * Periodic boundary conditions along the axes with a nonzero image count.
 * Replaces: the same set-up with pbc_images (ConvolutionDataCUDA::SetConvolutionDimensions,
 * Boris/ConvolutionDataCUDA.cpp:24-100; DemagKernelCUDA::Calculate_Demag_Kernels_2D/3D with
 * DemagTFunc::CalcDiagTens2D/3D_PBC and CalcOffDiagTens2D/3D_PBC,
 * Boris/DemagKernelCUDA.cpp:91-108/556-582, Boris/DemagTFunc_PBC.cpp:5-422), reached from
 * the `pbc` console command (Demag::Set_PBC).  Along a periodic axis N = n (no zero padding)
 * and the tensor sums the images -p..p of the source; in 2D, z images sit at whole multiples
 * of the cell thickness.  Sizes: N.x / 2 and N.y in the 2^a 3^b plans 2..4096 (so a periodic
 * nx is a multiple of 4), a periodic nz (3D) a power of two 4..256; anything else
 * returns BM_ERR_SIZE.  All three counts zero is bm_demag_create.  The convolution calls take
 * the handle unchanged; the integrator (bm_ode_set_pbc), the adjoint (bm_adj_set_pbc) and
 * batches (bm_batch_set_pbc) take periodic meshes; periodic layer stacks are made by
 * bm_mlayer_create_pbc.
 */
int bm_demag_create_pbc(int nx, int ny, int nz, double hx, double hy, double hz, int pbc_x, int pbc_y, int pbc_z,
	int flags, bm_demag_t* out);
int bm_demag_destroy(bm_demag_t d);
/* N.x, N.y, N.z of the zero-padded convolution and the device bytes held. */
int bm_demag_info(bm_demag_t d, int* N3, long long* device_bytes);

/*
 * Heff (=, +=) H_demag(M).
 * Replaces: ConvolutionCUDA<DemagCUDA,DemagKernelCUDA>::Convolute(M, Heff, energy,
 * get_energy, clearOut) (Boris/ConvolutionCUDA.h:217-223), called from
 * DemagCUDA::UpdateField (Boris/DemagCUDA.cpp:106-121).  clear_out = 1 sets Heff,
 * 0 adds to it.  If energy is non-NULL it receives (device float, overwritten)
 * -mu0/(2 n) * sum_cells M.H_demag, the reference's demag energy density.
 */
int bm_demag_convolute(bm_demag_t d, const float* M, float* Heff, int clear_out, float* energy, void* stream);
/* Averaged inputs, optionally duplicated outputs (two-sublattice meshes): H1 (=, +=) and, if
 * H2 is non-NULL, H2 (=, +=) the demag field of (M1 + M2) / 2; energy (device float or NULL)
 * = -mu0/2 <(M1 + M2)/2 . H> / n.  Replaces ConvolutionCUDA::Convolute_AveragedInputs and
 * Convolute_AveragedInputs_DuplicatedOutputs (Boris/ConvolutionCUDA.h:225-247), called by
 * DemagCUDA::UpdateField for antiferromagnetic meshes (Boris/DemagCUDA.cpp:125-129). */
int bm_demag_convolute_averaged(bm_demag_t d, const float* M1, const float* M2, float* H1, float* H2,
	int clear_out, float* energy, void* stream);
/* bm_demag_convolute_averaged without any device allocation: no counterpart in the
 * reference.  For callers that cannot allocate (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc): bm_demag_convolute_averaged allocates its
 * averaged-input and field buffers on first use, this variant keeps them in a static device
 * array that the library holds on each GPU until the process ends (24 MiB), so it serves
 * meshes of at most 1048576 cells (larger meshes return BM_ERR_SIZE).  Same arguments and
 * results (bit for bit) as bm_demag_convolute_averaged.  All calls on one GPU share that
 * GPU's array: calls that overlap on one GPU race. */
int bm_demag_convolute_averaged_static(bm_demag_t d, const float* M1, const float* M2, float* H1, float* H2,
	int clear_out, float* energy, void* stream);

/* H_demag(M) written as three planes Hx[n], Hy[n], Hz[n] (the layout the fused LLG
 * stage kernels read).  No reference counterpart: the split the reference makes between
 * Convolute and the field modules, with the output planar instead of interleaved. */
int bm_demag_convolute_planar(bm_demag_t d, const float* M, float* H3n, void* stream);

/* Copies the FP32 demag kernels to host memory (test hook).  3D: 6 planes Kxx, Kyy, Kzz,
 * Kxy, Kxz, Kyz each (N.x/2+1)*(N.y/2+1)*(N.z/2+1), index (kx*(N.y/2+1)+ky)*(N.z/2+1)+kz.
 * 2D: 4 planes Kxx, Kyy, Kzz, Kxy each (N.x/2+1)*(N.y/2+1), index kx*(N.y/2+1)+ky.
 * Values equal the reference's Kdiag/Kodiag/K2D_odiag (Boris/DemagKernel.cpp:171-516). */
int bm_demag_copy_kernels(bm_demag_t d, float* host_out, long long count);

/* ------------------------------------------------------------------------------------
 * Multilayered demag of stacked 2D layers: L (<= 64) layers of the same nx x ny cells
 * (hx, hy, hz[l]), layer l's bottom face at height z0[l] (m); layer thicknesses may differ.
 * M and H are L layers of nx*ny cells stacked (layer-major, AoS per cell).
 *
 * Replaces: SDemagCUDA's multilayered convolution for 2D layers related by z shifts —
 * SDemag::Initialize / set_Rect_collection (Boris/SDemag.cpp:173-300),
 * DemagKernelCollectionCUDA::Calculate_Demag_Kernels (Boris/DemagKernelCollectionCUDA_Calc.cpp:13;
 * self kernels :182, z-shifted :373; tensors Boris/DemagTFunc_Shifted.cpp:235-449 and
 * Boris/DemagTFunc_Irregular.cpp:5-235) computed once in FP64 on the GPU. */
typedef struct bm_mlayer_s* bm_mlayer_t;
int bm_mlayer_create(int nlayers, int nx, int ny, double hx, double hy, const double* hz, const double* z0,
	int flags, bm_mlayer_t* out);
/* The same stack with pbc_x, pbc_y periodic images along x and y (SDemag::Set_PBC,
 * Boris/SDemag.cpp:403-408; kernels DemagKernelCollectionCUDA_Calc.cpp:182/373 with pbc_images,
 * tensors Boris/DemagTFunc_Shifted_PBC.cpp:317-600 and DemagTFunc_Irregular_PBC.cpp:5-235):
 * N = n along a periodic axis (a periodic nx even with plans for nx / 2 and nx, a periodic ny
 * with a plan), no images along z.  Both counts zero is bm_mlayer_create.  The integrator
 * (bm_ode_set_layers, then bm_ode_set_pbc with the same x, y images) and the adjoint
 * (bm_adj_set_layers, bm_adj_set_pbc) take such a handle; batches do not. */
int bm_mlayer_create_pbc(int nlayers, int nx, int ny, double hx, double hy, const double* hz, const double* z0,
	int pbc_x, int pbc_y, int flags, bm_mlayer_t* out);
/*
 * Layers that are separate meshes of their own lateral size, cell and position: layer l has
 * nx x ny cells of hx x hy x hz, its bottom-left corner at (x0, y0, z0) (m).  Replaces the
 * same set-up over meshes of different extents and cells — SDemag::set_default_n_common /
 * set_Rect_collection (Boris/SDemag.cpp:173-300), which put every mesh's rectangle onto one
 * common grid, and the mesh transfer to and from it (SDemag_Demag::Initialize_Mesh_Transfer,
 * BorisLib/VEC_MeshTransfer.h).  The common cell is the reference's (per axis the largest
 * layer extent over the largest cell count), the transfer the reference's weights (in:
 * reciprocal distance times covered volume, out: reciprocal distance); the rectangles sit in
 * their bounding box at that cell (padded with empty cells to planned FFT lengths), and a
 * rectangle starting a fraction of a common cell off that grid couples to the others through
 * the tensor shifted in-plane by the fraction (complex kernels: the reference's Complex_Full
 * kernels, Boris/DemagKernelCollectionCUDA_Calc.cpp:571-745, tensors of the general branch of
 * Boris/DemagTFunc_Shifted.cpp:235-343 / DemagTFunc_Irregular.cpp:5-235).  M and H of all layers are the layers' own cells concatenated
 * layer by layer (AoS); the split calls take one layer's own nx*ny cells.  Such a handle
 * serves the convolution calls; the integrator takes handles of equal layers only. */
typedef struct { int nx, ny; double hx, hy, hz, x0, y0, z0; } bm_layer_mesh;
int bm_mlayer_create_meshes(int nlayers, const bm_layer_mesh* meshes, int flags, bm_mlayer_t* out);
int bm_mlayer_destroy(bm_mlayer_t m);
/* H (=, += if !clear_out) the demag field of every layer from the magnetisation of all
 * layers.  Replaces SDemagCUDA::UpdateField's multilayered branch (Boris/SDemagCUDA.cpp:270):
 * forward FFTs of all layers, KernelMultiplication_2D over all sources per destination
 * (Boris/DemagKernelCollectionCUDA_Mult.cu:610), inverse FFTs. */
int bm_mlayer_convolute(bm_mlayer_t m, const float* M, float* Heff, int clear_out, void* stream);
/* Split form, one call per layer mesh, for callers that hold each layer in its own array:
 * bm_mlayer_forward for every layer, then bm_mlayer_multiply once, then bm_mlayer_inverse
 * for every layer.  M_layer / H_layer are one layer's nx*ny cells (AoS); the inverse sets
 * (clear_out = 1) or adds to H_layer and, if energy (device float) is non-NULL, writes the
 * layer's -mu0/(2 nx ny) sum M.H_demag over its own cells into it (M_layer then needed).  Calls of one handle
 * share its spectra: one evaluation at a time.  Replaces the per-mesh
 * ConvolutionCUDA::ForwardFFT / KernelMultiplication_MultipleInputs / InverseFFT
 * (Boris/ConvolutionCUDA.h:254-351) of SDemagCUDA::UpdateField (Boris/SDemagCUDA.cpp:270+). */
int bm_mlayer_forward(bm_mlayer_t m, int layer, const float* M_layer, void* stream);
int bm_mlayer_multiply(bm_mlayer_t m, void* stream);
int bm_mlayer_inverse(bm_mlayer_t m, int layer, const float* M_layer, float* H_layer, int clear_out, float* energy, void* stream);
/* Copies the FP32 kernels to host memory (test hook): L*L pairs (destination-major), each
 * 6 planes Kxx, Kyy, Kzz, Kxy, Kxz, Kyz of (nx+1)*(ny+1), index kx*(ny+1)+ky.  Own-layer
 * pairs hold the 2D kernel (Kxz = Kyz = 0); Kxz and Kyz of z-shifted pairs are the
 * coefficients of i. */
int bm_mlayer_copy_kernels(bm_mlayer_t m, float* host_out, long long count);

/* ------------------------------------------------------------------------------------
 * LLG integrator for one ferromagnetic mesh: Zeeman (sets), demag, exchange and uniaxial
 * anisotropy (add) fused into each RK stage.
 *
 * Replaces: DifferentialEquationFMCUDA (Boris/DiffEqFMCUDA.h:38-137), its stage kernels
 * (Boris/DiffEqFM_EvalsCUDA_RKDP54.cu:15-330, Boris/DiffEqFM_EvalsLLGCUDA.cu:121-641),
 * the field modules' UpdateField (Boris/ZeemanCUDA.cu:215, Boris/ExchangeCUDA.cu:122,
 * Boris/AnisotropyUniCUDA.cu:112) and ODECommon_Base::IterateCUDA
 * (Boris/DiffEq_CommonBase_IterateCUDA.cpp:13) with SetAdaptiveTimeStep
 * (Boris/DiffEq_CommonBase_Iterate.cpp:15-59).
 */
typedef struct bm_ode_s* bm_ode_t;

/* evaluation methods: the reference's EVAL_ enum values (Boris/DiffEq_Defs.h) */
enum { BM_EVAL_TEULER = 1, BM_EVAL_RK4 = 2, BM_EVAL_AHEUN = 7, BM_EVAL_RKDP54 = 9 };

/* module switches */
enum { BM_MOD_DEMAG = 1, BM_MOD_EXCHANGE = 2, BM_MOD_ANIS_UNI = 4, BM_MOD_ZEEMAN = 8 };

typedef struct {
	double Ms;              /* A/m (used when Ms_map is NULL) */
	double A;               /* exchange stiffness J/m */
	double alpha;           /* Gilbert damping */
	double K1, K2;          /* uniaxial anisotropy J/m^3 */
	double ea[3];           /* easy axis (unit) */
	double grel;            /* gyromagnetic ratio scaling (1) */
	/* applied field H(t) = Ha + amp*dir*sinc(w*(t - t0)) [wave 1] or Ha + slope*t [wave 2],
	 * evaluated at the start of each step, A/m */
	double Ha[3];
	int wave;
	double amp, w, t0, dir[3], slope[3];
	int modules;            /* BM_MOD_* bits */
} bm_fm_params;

/* M: caller-owned device float[3n], the mesh magnetisation.  It holds the magnetisation
 * whenever a step is complete (after bm_ode_advance_time / bm_ode_run_to_iteration);
 * intermediate RK stages alternate between M and an internal array so that no stage
 * overwrites the magnetisation its exchange stencil is still reading.
 * demag may be NULL when BM_MOD_DEMAG is not set. */
int bm_ode_create(int method, int nx, int ny, int nz, double hx, double hy, double hz,
	const bm_fm_params* p, bm_demag_t demag, float* M, bm_ode_t* out);
int bm_ode_destroy(bm_ode_t o);
/* per-cell Ms map (device float[n], caller-owned, NULL to use p->Ms) */
int bm_ode_set_ms_map(bm_ode_t o, const float* Ms_map);
int bm_ode_set_params(bm_ode_t o, const bm_fm_params* p);
/* Layer stack: the mesh's nz planes are the layers of the multilayered demag ml (same nx,
 * ny, nz == layer count; at most 64), as the reference's one-ferromagnetic-mesh-per-layer
 * stack coupled by SDemag (Boris/SDemag.cpp:173-300, SDemagCUDA::UpdateField;
 * Boris/Exchange.cpp:108, Boris/Anisotropy.cpp:82 and Boris/DiffEqFM_Equations.cpp:12 per
 * layer mesh).  Exchange acts within each layer; Ms, A, alpha, K1 may be given per layer
 * (double[L] each, NULL keeps the mesh value; Ms becomes an internal per-cell map, replaced
 * by a later bm_ode_set_ms_map).  With BM_MOD_DEMAG set (bm_ode_set_params) the demag
 * field is the multilayered convolution.  ml must outlive o; ml NULL detaches the stack. */
int bm_ode_set_layers(bm_ode_t o, bm_mlayer_t ml, const double* Ms, const double* A, const double* alpha, const double* K1);
/*
 * Fields of modules outside the four fused ones (DMI, surface exchange, other anisotropies,
 * ...).  Before every stage the integrator calls fn(user, M, Hext, stream) with M the
 * magnetisation that stage evaluates (device float[3n], read only) and Hext the caller's
 * device float[3n]; fn writes the other modules' summed field into Hext (A/m), ordered on
 * `stream`, and the stage kernel adds it to the effective field.  fn NULL detaches it.
 * Replaces: the remaining modules' UpdateField calls of MeshBase::UpdateModulesCUDA
 * (Boris/MeshBaseModules.cpp:98-111, Boris/ModulesCUDA.h:114) that run before each
 * DifferentialEquationFMCUDA stage (Boris/DiffEqFMCUDA.h:38-112).  Single-mesh integrator
 * only (not the batch or adjoint handles).
 */
typedef void (*bm_field_fn)(void* user, const float* M, float* Hext, void* stream);
int bm_ode_set_extra_field(bm_ode_t o, bm_field_fn fn, void* user, float* Hext);
/* dT, and the adaptive controls err_fail, dT_increase, dT_min, dT_max (values <= 0 keep
 * the current ones); the defaults are the reference's per-method defaults. */
int bm_ode_set_dt(bm_ode_t o, double dT);
int bm_ode_set_astep(bm_ode_t o, double err_fail, double dT_increase, double dT_min, double dT_max);
/*
 * Stacks of 3D meshes (after bm_ode_set_layers): mesh_of_layer (nz ints, non-decreasing, or
 * NULL) marks consecutive layers that are the z planes of one ferromagnetic mesh, so that
 * mesh's exchange couples them (free ends at its top and bottom, spacing = the planes'
 * thickness; the planes must touch and be of equal thickness); Ha (3 doubles per layer, A/m,
 * or NULL) gives every layer its own constant applied field.
 * Replaces: the per-mesh modules of Boris's multilayered set-up of 3D meshes coupled by
 * SDemag (Boris/SDemag.cpp:173-300; each mesh's Exchange, Boris/Exchange.cpp:108, and
 * Zeeman, Boris/Zeeman.cpp:164).  The multilayered convolution of the planes, all pairs
 * z-shifted, gives the same interactions as the reference's 3D multilayered kernels.
 */
int bm_ode_set_layer_meshes(bm_ode_t o, const int* mesh_of_layer, const double* Ha);
/*
 * Periodic boundary conditions of the mesh (images per axis, 0 = open).
 * Replaces: the `pbc` command's mesh settings, Demag::Set_PBC + Mesh::Set_Magnetic_PBC
 * (Boris/Demag.cpp:137-148, Boris/MeshControl.cpp:49-65,
 * Boris/SuperMeshMeshes_Settings.cpp:659-700): exchange uses the cell on the other side as the
 * neighbour at the mesh ends of every axis with images (BorisLib/VEC_VC_Del.h:10-130); with
 * demag on, the handle given at creation must have been made by bm_demag_create_pbc with the
 * same images (BM_ERR_STATE otherwise), as the reference sets both together.  The evaluation
 * speedup's self coefficient becomes DemagTFunc::SelfDemag_PBC with these images
 * (Boris/DemagTFunc_Self.cpp:58-99).  Not for layer stacks.
 */
int bm_ode_set_pbc(bm_ode_t o, int pbc_x, int pbc_y, int pbc_z);
/* Demag evaluation speedup: mode is the reference's EVALSPEEDUP_ value (0 none, 1 step,
 * 2 linear, 3 quadratic, 4 cubic, 5 quartic, 6 quintic; Boris/DiffEq_Defs.h:121-125) =
 * the number of saved demag fields.  The demag field is convolved only when the ODE
 * recommends an update (at the first stage of each step, or every dTspeedup of simulated
 * time when dTspeedup > 0) and extrapolated in time from the saved fields otherwise.
 * Replaces ODECommon_Base::SetEvaluationSpeedup / SetSpeedupTimeStep / SetLink_dTspeedup
 * (Boris/DiffEq_CommonBase.h:252, Boris/DiffEq_CommonBase_Control.cpp:243-256) with
 * DemagCUDA::UpdateField's speedup branch (Boris/DemagCUDA.cpp:106-468); for a layer stack
 * (bm_ode_set_layers) SDemagCUDA::UpdateField's multilayered speedup branch
 * (Boris/SDemagCUDA.cpp:583-1000), each layer with its own self-demag coefficient.
 * Allocates one field array per saved field; drops saved fields. */
int bm_ode_set_eval_speedup(bm_ode_t o, int mode, double dTspeedup);
/* Saved-field count, their save times (double[6]) and the number of convolutions run so
 * far (any pointer may be NULL).  DemagCUDA::num_Hdemag_saved / time_demag1..6
 * (Boris/DemagCUDA.h:35-37). */
int bm_ode_get_speedup_state(bm_ode_t o, int* num_saved, double* tdm6, long long* convolutions);

/* One SuperMesh::AdvanceTimeCUDA call (Boris/SuperMeshSimulation.cpp:122-149): field
 * updates and stages until the step is solved (adaptive methods may reject and redo). */
int bm_ode_advance_time(bm_ode_t o, void* stream);
/* Advance until the accepted-iteration counter reaches `iteration`. */
int bm_ode_run_to_iteration(bm_ode_t o, long long iteration, void* stream);
/* One field update + one stage (ODECommon_Base::IterateCUDA granularity). */
int bm_ode_iterate_stage(bm_ode_t o, void* stream);
/* M <- sM1 (DifferentialEquationFMCUDA::RestoreMagnetization). */
int bm_ode_restore(bm_ode_t o, void* stream);

typedef struct {
	double time, stagetime, dT, dT_last, lte;
	long long iteration, stageiteration, rejections, evaluations;
	int evalStep, primed, available;
} bm_ode_state;
int bm_ode_get_state(bm_ode_t o, bm_ode_state* st);
/* Keep a copy of the effective field of each stage (off by default: it costs one more
 * array write per stage).  Test hook; no reference counterpart. */
int bm_ode_keep_heff(bm_ode_t o, int on);
/* Effective field of the last stage (device float[3n], caller-owned); needs keep_heff. */
int bm_ode_copy_heff(bm_ode_t o, float* Heff_out, void* stream);
/* <M>/Ms averaged over cells (host double[3]); synchronises the stream. */
int bm_ode_average_m(bm_ode_t o, double* out3, void* stream);

/* ------------------------------------------------------------------------------------
 * Batched integration of many independent meshes of uneven sizes (no reference
 * counterpart: the reference integrates one mesh set per simulation).  Mesh b has
 * n3[3b..3b+2] cells of h3[3b..3b+2] (m) and its own parameters params[b]; M is the
 * caller's device array holding every mesh's magnetisation concatenated in mesh order
 * (bm_batch_mesh_offset gives each mesh's cell offset).  Methods: BM_EVAL_RK4,
 * BM_EVAL_TEULER (fixed step) and BM_EVAL_RKDP54 (every mesh with its own adaptive step,
 * controlled on the device as ODECommon_Base::Iterate does for one mesh,
 * Boris/DiffEq_CommonBase_Iterate.cpp:15-59,1214-1395); applied fields constant, sinc
 * pulse or ramp, evaluated at each step's start time on the device.  Every pass of a field evaluation is one launch per
 * distinct pass size over all meshes of that size; a whole step is captured once into a
 * hipGraph and replayed (BM_BATCH_NO_GRAPH: launch directly).  BM_BATCH_AUTOTUNE: before
 * the first step, time candidate launch shapes of every pass-size group (x-pass row-tile
 * factor, workgroups per CU) on the real tables and keep the fastest;
 * BM_BATCH_AUTOTUNE_RADIX (implies BM_BATCH_AUTOTUNE): the y, z and 2D y passes' radix
 * schedules (alternative plans of each FFT length) are timed as well; results then differ
 * from the default schedules by FP32 rounding (launch shapes alone change no bit).
 * bm_batch_autotune_report writes the choices and times as text lines
 * ("<pass> <size> f=<f> per_cu=<n> radix=<v> <radices> <us> us"); bm_batch_set_schedule
 * pins one pass-size group's choice (e.g. read back from an earlier report), replacing any
 * pending tuning: f (x passes: 0 = default, 1 the base row tile, 2 twice it, -2 half of it
 * where that holds a whole row group: nx 128 to 768; other passes 0), per_cu (0 = default, <= 64),
 * radix (0 = default; y/z/y2d passes only); BM_ERR_ARG for a pass or size the batch lacks.
 * bm_batch_set_pbc (before the first bm_batch_run): mesh `mesh` becomes periodic as
 * bm_ode_set_pbc makes one mesh (its demag handle rebuilt as bm_demag_create_pbc with these
 * images, exchange wrapping around; not for layer stacks); its passes along periodic axes
 * form groups of their own, reported and pinned with "_pbc" appended to the pass name
 * (default radix schedule, x-pass factor 1); pass groups are rebuilt, so set the images before
 * pinning schedules.
 * bm_batch_create_layers: layers[b].nlayers > 0 makes mesh b a layer stack as
 * bm_ode_set_layers does (nz == nlayers <= 16, demag on; hz, z0 its layers' thicknesses and
 * bottom heights, Ms/A/alpha/K1 per layer or NULL); mesh_of_layer and Ha (or NULL) make it a
 * stack of 3D meshes as bm_ode_set_layer_meshes does; layers NULL = bm_batch_create. */
typedef struct bm_batch_s* bm_batch_t;
enum { BM_BATCH_NO_GRAPH = 1, BM_BATCH_AUTOTUNE = 2, BM_BATCH_AUTOTUNE_RADIX = 4 };
int bm_batch_create(int method, int nmesh, const int* n3, const double* h3, const bm_fm_params* params,
	float* M, int flags, bm_batch_t* out);
typedef struct {
	int nlayers;                                    /* 0: a plain mesh */
	const double* hz; const double* z0;             /* nlayers values each (m) */
	const double* Ms; const double* A; const double* alpha; const double* K1;   /* per layer or NULL */
	const int* mesh_of_layer;                       /* stacks of 3D meshes: mesh index per layer, or NULL */
	const double* Ha;                               /* constant applied field per layer (3 each, A/m), or NULL */
} bm_layer_desc;
int bm_batch_create_layers(int method, int nmesh, const int* n3, const double* h3, const bm_fm_params* params,
	const bm_layer_desc* layers, float* M, int flags, bm_batch_t* out);
int bm_batch_destroy(bm_batch_t b);
int bm_batch_set_dt(bm_batch_t b, double dT);
/* nsteps steps of every mesh (RKDP54: nsteps accepted steps of each mesh; rejected attempts
 * are redone), ordered after earlier work on `stream`. */
int bm_batch_run(bm_batch_t b, long long nsteps, void* stream);
int bm_batch_get_state(bm_batch_t b, double* time, long long* iteration);
/* RKDP54 batches: the adaptive controls err_fail, dT_increase, dT_min, dT_max of every mesh
 * (values <= 0 keep the current ones; defaults are the reference's RKDP54 defaults).
 * Replaces ODECommon_Base::SetAdaptiveTimeStepCtrl (Boris/DiffEq_CommonBase_Control.cpp:258). */
int bm_batch_set_astep(bm_batch_t b, double err_fail, double dT_increase, double dT_min, double dT_max);
/* One mesh's time, dT, accepted iterations and rejected attempts after the last run
 * (fixed-step batches: the common values, no rejections). */
int bm_batch_get_mesh_state(bm_batch_t b, int mesh, double* time, double* dT, long long* iteration, long long* rejections);
int bm_batch_mesh_offset(bm_batch_t b, int mesh, long long* cell_offset, long long* ncells);
int bm_batch_autotune_report(bm_batch_t b, char* out, int cap);
int bm_batch_set_schedule(bm_batch_t b, const char* pass, int size, int f, int per_cu, int radix);
int bm_batch_set_pbc(bm_batch_t b, int mesh, int pbc_x, int pbc_y, int pbc_z);

/* ------------------------------------------------------------------------------------
 * Differentiation of the fixed-step RK4 rollout (no reference counterpart: the reference
 * has no differentiation).  Forward: nsteps RK4 steps of dt from M0_i = Ms_i m0_i with
 * demag (bm_demag handle of this mesh), exchange, uniaxial anisotropy and Zeeman
 * H(t) = Ha + amp dir sinc(w (t - t0)) (params->wave = 1) or Ha + slope t (wave = 2), the
 * per-cell Ms map Ms_map (device, caller-owned, n floats); loss
 * L = 1 - (1/n) sum_i (M_i/Ms_i) . mt_i.  Backward: the discrete adjoint, dL/dMs (device,
 * n floats) and dL/d amp; bm_adj_zeeman_grads then gives (dL/d amp, dL/d slope[3]).  The
 * state is checkpointed every ckpt_every steps and each segment recomputed once in the
 * backward pass.  bm_adj_set_layers (before bm_adj_forward) makes the nz planes a layer
 * stack as bm_ode_set_layers does (multilayered demag with its exact transpose in the
 * backward pass, per-layer A, alpha, K1; per-layer Ms through Ms_map).  bm_adj_set_pbc
 * (before bm_adj_forward) makes the mesh periodic as bm_ode_set_pbc does: exchange wraps
 * around along each axis with images and, with demag on, the handle must have been made by
 * bm_demag_create_pbc with the same images (BM_ERR_STATE otherwise); not for layer stacks. */
typedef struct bm_adj_s* bm_adj_t;
int bm_adj_create(int nx, int ny, int nz, double hx, double hy, double hz, const bm_fm_params* params,
	bm_demag_t demag, const float* Ms_map, double dt, int nsteps, int ckpt_every, bm_adj_t* out);
int bm_adj_destroy(bm_adj_t a);
/* m0 and mt: device float[3n] unit vectors; M_final (optional, device float[3n]). */
int bm_adj_forward(bm_adj_t a, const float* m0, const float* mt, float* M_final, double* loss, void* stream);
int bm_adj_backward(bm_adj_t a, float* dL_dMs, double* dL_damp, void* stream);
int bm_adj_zeeman_grads(bm_adj_t a, double* out4);
int bm_adj_set_layers(bm_adj_t a, bm_mlayer_t ml, const double* A, const double* alpha, const double* K1);
int bm_adj_set_pbc(bm_adj_t a, int pbc_x, int pbc_y, int pbc_z);
/* Batched adjoint (no reference counterpart): `count` adjoint handles driven in lockstep, each
 * on its own internal stream so that the instances' kernels overlap on the device (the
 * differentiable multilayer workload runs 16 stacks this way).  Work is ordered after earlier
 * work on `stream`, which waits for all of it on return.  All handles need the same nsteps and
 * checkpoint interval and their own demag / multilayer handles.  m0[b], mt[b], M_final[b] (or
 * NULL), gMs_out[b] per handle; loss and gamp are double[count]; bm_adj_zeeman_grads of each
 * handle gives its field-waveform gradients afterwards. */
int bm_adj_batch_forward(int count, const bm_adj_t* h, const float* const* m0, const float* const* mt,
	float* const* M_final, double* loss, void* stream);
int bm_adj_batch_backward(int count, const bm_adj_t* h, float* const* gMs_out, double* gamp, void* stream);

/* ------------------------------------------------------------------------------------
 * Launch timing (no reference counterpart: the reference only times whole runs,
 * Boris/Commands.cpp:4759).  Between bm_prof_begin and bm_prof_end every hot-path kernel
 * launch is bracketed by HIP events recorded on the launch's own stream; bm_prof_query
 * returns the count and the total / min / max duration (ms) of one kind:
 * 0 x-R2C pass, 1 y-FFT pass, 2 z-pass with kernel multiply, 3 y-IFFT pass,
 * 4 x-C2R pass, 5 2D y pass with kernel multiply, 6 field + RK stage (no demag),
 * 7 RKDP54 advance (no demag), 8 fused row kernel (x-C2R of H_demag + field + RK stage +
 * x-R2C of the stage's output M), 9 x-R2C pass fused with the RKDP54 advance,
 * 10 multilayer kernel product.
 * Not thread safe; meant for benchmarks. */
int bm_prof_begin(int max_launches);
int bm_prof_end(void);
int bm_prof_query(int kind, long long* count, double* total_ms, double* min_ms, double* max_ms);

/* ------------------------------------------------------------------------------------
 * OVF 2.0 files (host memory, no device work).  Replaces OVF2::Read_OVF2_VEC /
 * Read_OVF2_SCA / Write_OVF2_VEC / Write_OVF2_SCA (Boris/OVF2_Handlers.cpp:41-902), used by
 * the reference's loadovf2mag / saveovf2mag commands for magnetisation states.  Rectangular
 * meshes, meshunit m or nm, valuedim 1 or 3, data "binary 4", "binary 8" or "text", x
 * fastest; sizes in metres.
 */
/* Header: node counts n3, cell size h3, rectangle rect6 = (xmin, ymin, zmin, xmax, ymax,
 * zmax), valuedim (1 or 3) and data_bytes (4, 8, or 1 for text); any pointer may be NULL. */
int bm_ovf2_info(const char* path, int* n3, double* h3, double* rect6, int* valuedim, int* data_bytes);
/* The data as doubles: count = valuedim * cells values, cell-major, x fastest. */
int bm_ovf2_read(const char* path, double* out, long long count);
/* Writes valuedim * cells values (divided by norm) with the reference's header layout;
 * data_type 4 = "binary 4", 8 = "binary 8", 0 = "text".  The title is the file name. */
int bm_ovf2_write(const char* path, const double* data, int valuedim, const int* n3, const double* h3,
	const double* rect6, int data_type, double norm);

#ifdef __cplusplus
}
#endif

#endif
