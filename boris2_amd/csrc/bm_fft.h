// Device FFT engine for CDNA4 (gfx950), FP32 and FP64.
//
// One transform of length N is carried by T = N/E threads, each holding E complex values
// in registers.  Stages are Stockham radix-R passes (R divides E); between stages values
// go through LDS once (one write + one read per element), the first stage consumes the
// registers directly and the last stage produces them directly, so a pass kernel can load
// from and store to HBM in natural order with coalesced accesses:
//     on entry and on exit, register v[e] of thread t holds element  t + T*e.
// Zero padding: when only the first N/2 inputs are non-zero, callers set v[e] = 0 for
// e >= E/2 and the compiler folds the zeros through the first radix butterfly; when only
// the first N/2 outputs are needed, callers store only e < E/2 and the dead outputs of the
// last butterfly are eliminated.
//
// Twiddles come from FP64 sincospi (rounded to the working precision), computed once per
// thread per kernel (they depend only on the thread's butterfly indices) and kept in
// registers across every line the thread transforms.
#pragma once

#include <hip/hip_runtime.h>

namespace bm {

template <typename T> struct C2 { T x, y; };

template <typename T> __device__ __forceinline__ C2<T> mk(T a, T b) { C2<T> r; r.x = a; r.y = b; return r; }
template <typename T> __device__ __forceinline__ C2<T> operator+(C2<T> a, C2<T> b) { return mk(a.x + b.x, a.y + b.y); }
template <typename T> __device__ __forceinline__ C2<T> operator-(C2<T> a, C2<T> b) { return mk(a.x - b.x, a.y - b.y); }
template <typename T> __device__ __forceinline__ C2<T> operator*(T s, C2<T> a) { return mk(s * a.x, s * a.y); }
template <typename T> __device__ __forceinline__ C2<T> cmul(C2<T> a, C2<T> b) { return mk(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
// a * conj(b)
template <typename T> __device__ __forceinline__ C2<T> cmulc(C2<T> a, C2<T> b) { return mk(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y); }
template <typename T> __device__ __forceinline__ C2<T> conj(C2<T> a) { return mk(a.x, -a.y); }
// multiply by DIR*i  (DIR = -1 forward, +1 inverse)
template <int DIR, typename T> __device__ __forceinline__ C2<T> mul_di(C2<T> a) { return DIR < 0 ? mk(a.y, -a.x) : mk(-a.y, a.x); }

// ----------------------------------------------------------------------------------------
// Radix plans.  E = values per thread; each radix divides E; the first radix is even so a
// zero upper half of the input maps onto whole zero inputs of the first butterflies.
// Workgroups are dispatched round-robin over the 8 XCDs (each with its own L2).  Item
// c = (b mod 8) * ceil(nb/8) + b/8 of a sweep of nb items goes to block b, so each XCD works
// on one contiguous eighth of the sweep and neighbouring items share an L2.  nb must be a
// multiple of 8 for the map to be a permutation of 0..nb-1.
__device__ __forceinline__ int xcd_chunk(int b, int nb)
{
	const int per = (nb + 7) / 8;
	return (b % 8) * per + b / 8;
}

// A plan is looked up by its key: the length N for the default schedule, N + 65536 v for
// alternative schedule v >= 1 of the same length (other E and radices; the batched
// autotuner times them against the default).  Template parameters named N below are keys.
template <int K> struct Plan;
constexpr int plan_key(int n, int v) { return n + 65536 * v; }
#define BM_PLAN(n, e, ...) BM_PLANV(n, 0, e, __VA_ARGS__)
#define BM_PLANV(n, v, e, ...)                                                  \
	template <> struct Plan<plan_key(n, v)> {                                   \
		static constexpr int L = n;                                             \
		static constexpr int E = e;                                             \
		static constexpr int R[] = { __VA_ARGS__ };                             \
		static constexpr int S = sizeof(R) / sizeof(int);                       \
	};
BM_PLAN(2, 2, 2)
BM_PLAN(6, 6, 2, 3)
BM_PLAN(12, 12, 4, 3)
BM_PLAN(24, 12, 4, 3, 2)
BM_PLAN(48, 12, 4, 4, 3)
BM_PLAN(4, 4, 4)
BM_PLAN(8, 8, 8)
BM_PLAN(16, 8, 8, 2)
BM_PLAN(32, 8, 8, 4)
BM_PLAN(64, 8, 8, 8)
BM_PLAN(96, 12, 4, 4, 3, 2)
BM_PLAN(128, 8, 8, 4, 4)
BM_PLAN(192, 12, 4, 4, 4, 3)
BM_PLAN(256, 8, 8, 8, 4)
BM_PLAN(384, 12, 4, 4, 4, 3, 2)
BM_PLAN(512, 8, 8, 8, 8)
BM_PLAN(768, 12, 4, 4, 4, 4, 3)
BM_PLAN(1024, 16, 16, 16, 4)
BM_PLAN(1536, 12, 4, 4, 4, 4, 3, 2)
BM_PLAN(2048, 16, 16, 16, 8)
BM_PLAN(4096, 16, 16, 16, 16)
// alternative schedules: v = 1 fewer threads per line (larger radices, fewer LDS exchanges),
// v = 2 more threads per line (smaller radices)
BM_PLANV(4, 1, 2, 2, 2)
BM_PLANV(8, 1, 4, 4, 2)
BM_PLANV(8, 2, 2, 2, 2, 2)
BM_PLANV(16, 1, 16, 16)
BM_PLANV(16, 2, 4, 4, 4)
BM_PLANV(32, 1, 16, 16, 2)
BM_PLANV(32, 2, 4, 4, 4, 2)
BM_PLANV(64, 1, 16, 16, 4)
BM_PLANV(64, 2, 4, 4, 4, 4)
BM_PLANV(96, 1, 24, 8, 4, 3)
BM_PLANV(128, 1, 16, 16, 8)
BM_PLANV(128, 2, 4, 4, 4, 4, 2)
BM_PLANV(192, 1, 24, 8, 8, 3)
BM_PLANV(256, 1, 16, 16, 16)
BM_PLANV(256, 2, 4, 4, 4, 4, 4)
BM_PLANV(384, 1, 24, 8, 8, 3, 2)
BM_PLANV(512, 1, 16, 16, 8, 4)
BM_PLANV(512, 2, 4, 4, 4, 4, 4, 2)
BM_PLANV(768, 1, 24, 8, 8, 4, 3)
BM_PLANV(1024, 1, 8, 8, 8, 8, 2)
BM_PLANV(1024, 2, 4, 4, 4, 4, 4, 4)
BM_PLANV(2048, 1, 8, 8, 8, 8, 4)
#undef BM_PLAN
#undef BM_PLANV
// number of schedules (the default and its alternatives) of length n
constexpr int plan_variants(int n)
{
	switch (n) {
	case 4: case 96: case 192: case 384: case 768: case 2048: return 2;
	case 8: case 16: case 32: case 64: case 128: case 256: case 512: case 1024: return 3;
	default: return 1;
	}
}

template <int N> constexpr int plan_T() { return Plan<N>::L / Plan<N>::E; }
template <int N> constexpr int plan_Ns(int s) { int p = 1; for (int q = 0; q < s; q++) p *= Plan<N>::R[q]; return p; }
template <int N> constexpr int plan_ntw()
{
	int n = 0;
	for (int s = 1; s < Plan<N>::S; s++) n += (Plan<N>::E / Plan<N>::R[s]) * (Plan<N>::R[s] - 1);
	return n > 0 ? n : 1;
}
template <int N> constexpr int plan_twoff(int s)
{
	int n = 0;
	for (int q = 1; q < s; q++) n += (Plan<N>::E / Plan<N>::R[q]) * (Plan<N>::R[q] - 1);
	return n;
}

// ----------------------------------------------------------------------------------------
// In-register DFTs of length R, direction DIR (-1: exp(-2 pi i jk/R)).
template <int R, int DIR, typename T> struct Dft;

template <int DIR, typename T> struct Dft<2, DIR, T> {
	static __device__ __forceinline__ void run(C2<T>* v) { C2<T> a = v[0], b = v[1]; v[0] = a + b; v[1] = a - b; }
};

template <int DIR, typename T> struct Dft<3, DIR, T> {
	static __device__ __forceinline__ void run(C2<T>* v)
	{
		const T c = (T)-0.5, s = (T)(DIR * 0.86602540378443864676372317075294);
		C2<T> t1 = v[1] + v[2], t2 = v[1] - v[2];
		C2<T> m = v[0] + c * t1;
		C2<T> r = mk(-s * t2.y, s * t2.x);   // i*s*t2
		v[0] = v[0] + t1; v[1] = m + r; v[2] = m - r;
	}
};

template <int DIR, typename T> struct Dft<4, DIR, T> {
	static __device__ __forceinline__ void run(C2<T>* v)
	{
		C2<T> a0 = v[0] + v[2], a1 = v[0] - v[2], a2 = v[1] + v[3], a3 = mul_di<DIR>(v[1] - v[3]);
		v[0] = a0 + a2; v[1] = a1 + a3; v[2] = a0 - a2; v[3] = a1 - a3;
	}
};

template <int DIR, typename T> struct Dft<8, DIR, T> {
	static __device__ __forceinline__ void run(C2<T>* v)
	{
		C2<T> e[4] = { v[0], v[2], v[4], v[6] }, o[4] = { v[1], v[3], v[5], v[7] };
		Dft<4, DIR, T>::run(e); Dft<4, DIR, T>::run(o);
		const T r = (T)0.70710678118654752440084436210485;
		// o[k] *= w8^k, w8 = exp(DIR i pi/4)
		o[1] = mk(r * (o[1].x - DIR * o[1].y), r * (o[1].y + DIR * o[1].x));
		o[2] = mul_di<DIR>(o[2]);
		o[3] = mk(r * (-o[3].x - DIR * o[3].y), r * (-o[3].y + DIR * o[3].x));
		for (int k = 0; k < 4; k++) { v[k] = e[k] + o[k]; v[k + 4] = e[k] - o[k]; }
	}
};

template <int DIR, typename T> struct Dft<16, DIR, T> {
	static __device__ __forceinline__ void run(C2<T>* v)
	{
		// 4 x 4: columns q (stride 4), twiddle w16^(q k1), then rows.
		C2<T> a[4][4];
#pragma unroll
		for (int q = 0; q < 4; q++) {
			C2<T> c[4] = { v[q], v[q + 4], v[q + 8], v[q + 12] };
			Dft<4, DIR, T>::run(c);
#pragma unroll
			for (int k = 0; k < 4; k++) a[q][k] = c[k];
		}
		const T c1 = (T)0.92387953251128675612818318939679, s1 = (T)0.38268343236508977172845998403040;
		const T r2 = (T)0.70710678118654752440084436210485;
		// w16^m = exp(DIR i 2 pi m / 16) for m = q*k1 in {1,2,3,4,6,9}
		auto tw = [&](C2<T> x, T c, T s) { s = (T)DIR * s; return mk(x.x * c - x.y * s, x.x * s + x.y * c); };
		a[1][1] = tw(a[1][1], c1, s1); a[1][2] = tw(a[1][2], r2, r2); a[1][3] = tw(a[1][3], s1, c1);
		a[2][1] = tw(a[2][1], r2, r2); a[2][2] = mul_di<DIR>(a[2][2]); a[2][3] = tw(a[2][3], -r2, r2);
		a[3][1] = tw(a[3][1], s1, c1); a[3][2] = tw(a[3][2], -r2, r2); a[3][3] = tw(a[3][3], -c1, -s1);
#pragma unroll
		for (int k = 0; k < 4; k++) {
			C2<T> c[4] = { a[0][k], a[1][k], a[2][k], a[3][k] };
			Dft<4, DIR, T>::run(c);
#pragma unroll
			for (int m = 0; m < 4; m++) v[k + 4 * m] = c[m];
		}
	}
};

// ----------------------------------------------------------------------------------------
// Per-thread twiddle cache: forward twiddles exp(-2 pi i (j mod Ns) r / (Ns R)) for every
// stage s >= 1, butterfly b and r >= 1.  The inverse direction uses their conjugates.
template <int N, typename T> struct Twiddles {
	C2<T> w[plan_ntw<N>()];

	template <int s>
	__device__ __forceinline__ void init_stage(int t)
	{
		if constexpr (s < Plan<N>::S) {
			constexpr int E = Plan<N>::E, TT = plan_T<N>(), R = Plan<N>::R[s], B = E / R, Ns = plan_Ns<N>(s);
			constexpr int off = plan_twoff<N>(s);
#pragma unroll
			for (int b = 0; b < B; b++) {
				int j = t + b * TT, k = j % Ns;
#pragma unroll
				for (int r = 1; r < R; r++) {
					double sn, cs;
					sincospi(-2.0 * (double)(k * r) / (double)(Ns * R), &sn, &cs);
					w[off + b * (R - 1) + r - 1] = mk((T)cs, (T)sn);
				}
			}
			init_stage<s + 1>(t);
		}
	}

	__device__ __forceinline__ void init(int t) { init_stage<1>(t); }
};

// Barrier of the LDS exchanges between FFT stages.  WAVE: every line of the exchange is held
// by lanes of one wavefront (T <= 64 threads per line, lines never straddling two waves), so
// ordering the wave's own LDS accesses suffices: a wave's LDS instructions execute in issue
// order, and the wavefront-scope fences keep the compiler from moving LDS accesses across
// the point.  Otherwise the whole workgroup waits.
template <bool WAVE>
__device__ __forceinline__ void xsync()
{
	if constexpr (WAVE) {
		__builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
		__builtin_amdgcn_wave_barrier();
		__builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
	}
	else __syncthreads();
}

// LDS element i of a line: lines with unit stride are padded by one element every 16 so
// the radix-R scatter of a Stockham stage (lane stride R elements) hits distinct banks;
// strided lines (several lines interleaved element by element) are conflict-free as is.
__device__ __forceinline__ int lidx(int i, int stride) { return stride == 1 ? i + (i >> 4) : i * stride; }
// LDS elements one padded unit-stride line of length N occupies
template <int N> constexpr int lds_len() { return N + N / 16 + 1; }

// One Stockham stage on registers.  LDS element i of this line lives at lds[lidx(i, stride)].
template <int N, int s, int DIR, typename T, bool WAVE = false>
__device__ __forceinline__ void fft_stage(C2<T>* v, C2<T>* lds, int stride, int t, const Twiddles<N, T>& tw)
{
	constexpr int E = Plan<N>::E, TT = plan_T<N>(), R = Plan<N>::R[s], B = E / R, Ns = plan_Ns<N>(s);
	if constexpr (s > 0) {
#pragma unroll
		for (int b = 0; b < B; b++)
#pragma unroll
			for (int r = 0; r < R; r++) v[b + r * B] = lds[lidx(t + b * TT + r * (Plan<N>::L / R), stride)];
#pragma unroll
		for (int b = 0; b < B; b++)
#pragma unroll
			for (int r = 1; r < R; r++) {
				C2<T> w = tw.w[plan_twoff<N>(s) + b * (R - 1) + r - 1];
				v[b + r * B] = (DIR < 0) ? cmul(v[b + r * B], w) : cmulc(v[b + r * B], w);
			}
	}
#pragma unroll
	for (int b = 0; b < B; b++) {
		C2<T> u[R];
#pragma unroll
		for (int r = 0; r < R; r++) u[r] = v[b + r * B];
		Dft<R, DIR, T>::run(u);
#pragma unroll
		for (int r = 0; r < R; r++) v[b + r * B] = u[r];
	}
	if constexpr (s < Plan<N>::S - 1) {
		xsync<WAVE>();
#pragma unroll
		for (int b = 0; b < B; b++) {
			int j = t + b * TT;
			int d = (j / Ns) * Ns * R + (j % Ns);
#pragma unroll
			for (int r = 0; r < R; r++) lds[lidx(d + r * Ns, stride)] = v[b + r * B];
		}
		xsync<WAVE>();
	}
}

template <int N, int s, int DIR, typename T, bool WAVE = false>
__device__ __forceinline__ void fft_stages(C2<T>* v, C2<T>* lds, int stride, int t, const Twiddles<N, T>& tw)
{
	fft_stage<N, s, DIR, T, WAVE>(v, lds, stride, t, tw);
	if constexpr (s + 1 < Plan<N>::S) fft_stages<N, s + 1, DIR, T, WAVE>(v, lds, stride, t, tw);
}

// Full transform.  All threads of the block (WAVE: of the wavefront) must call it together
// (it synchronises between stages); lines handled by different thread groups must use
// disjoint LDS regions.
template <int N, int DIR, typename T, bool WAVE = false>
__device__ __forceinline__ void fft(C2<T>* v, C2<T>* lds, int stride, int t, const Twiddles<N, T>& tw)
{
	fft_stages<N, 0, DIR, T, WAVE>(v, lds, stride, t, tw);
}

// Three transforms of the same length in lock step (shared barriers).
template <int N, int s, int DIR, typename T>
__device__ __forceinline__ void fft3_stages(C2<T>* v0, C2<T>* v1, C2<T>* v2, C2<T>* l0, C2<T>* l1, C2<T>* l2,
	int stride, int t, const Twiddles<N, T>& tw)
{
	constexpr int E = Plan<N>::E, TT = plan_T<N>(), R = Plan<N>::R[s], B = E / R, Ns = plan_Ns<N>(s);
	C2<T>* vv[3] = { v0, v1, v2 };
	C2<T>* ll[3] = { l0, l1, l2 };
	if constexpr (s > 0) {
#pragma unroll
		for (int c = 0; c < 3; c++) {
#pragma unroll
			for (int b = 0; b < B; b++)
#pragma unroll
				for (int r = 0; r < R; r++) vv[c][b + r * B] = ll[c][lidx(t + b * TT + r * (Plan<N>::L / R), stride)];
#pragma unroll
			for (int b = 0; b < B; b++)
#pragma unroll
				for (int r = 1; r < R; r++) {
					C2<T> w = tw.w[plan_twoff<N>(s) + b * (R - 1) + r - 1];
					vv[c][b + r * B] = (DIR < 0) ? cmul(vv[c][b + r * B], w) : cmulc(vv[c][b + r * B], w);
				}
		}
	}
#pragma unroll
	for (int c = 0; c < 3; c++)
#pragma unroll
		for (int b = 0; b < B; b++) {
			C2<T> u[R];
#pragma unroll
			for (int r = 0; r < R; r++) u[r] = vv[c][b + r * B];
			Dft<R, DIR, T>::run(u);
#pragma unroll
			for (int r = 0; r < R; r++) vv[c][b + r * B] = u[r];
		}
	if constexpr (s < Plan<N>::S - 1) {
		__syncthreads();
#pragma unroll
		for (int c = 0; c < 3; c++)
#pragma unroll
			for (int b = 0; b < B; b++) {
				int j = t + b * TT;
				int d = (j / Ns) * Ns * R + (j % Ns);
#pragma unroll
				for (int r = 0; r < R; r++) ll[c][lidx(d + r * Ns, stride)] = vv[c][b + r * B];
			}
		__syncthreads();
		fft3_stages<N, s + 1, DIR, T>(v0, v1, v2, l0, l1, l2, stride, t, tw);
	}
}

template <int N, int DIR, typename T>
__device__ __forceinline__ void fft3(C2<T>* v0, C2<T>* v1, C2<T>* v2, C2<T>* l0, C2<T>* l1, C2<T>* l2,
	int stride, int t, const Twiddles<N, T>& tw)
{
	fft3_stages<N, 0, DIR, T>(v0, v1, v2, l0, l1, l2, stride, t, tw);
}

} // namespace bm
