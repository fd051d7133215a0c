// Learned transfer-function surrogate: a 1-32-32-1 SiLU MLP per transfer
// function, trained on the GPU by the kernels below on samples the exact
// function generates, evaluated on matrix cores, and used only where a
// per-bin validation against the exact function stays inside a tolerance;
// everywhere else the exact transfer (the reference's GammaOperationC
// arithmetic, operation_impl.cpp:53-77) is evaluated instead.
//
// Network: z1 = w1 u + b1, h1 = silu(z1), z2 = W2 h1 + b2, h2 = silu(z2),
// v = w3 . h2 + b3, silu(z) = z / (1 + e^-z). u is the input mapped to
// [-1, 1] (optionally through log2), v the normalised output (optionally
// mapped back through exp2). The 32 x 32 layer (and, in training, its two
// gradient products) runs on v_mfma_f32_16x16x32_f16 with FP16 high + low
// split operands and FP32 accumulation; the rank-1 layers stay in FP32.
#include "learned.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>

#include "common.h"
#include "zmath.h"

namespace zamd {

namespace {

typedef _Float16 h8 __attribute__((ext_vector_type(8)));
typedef float f4 __attribute__((ext_vector_type(4)));

constexpr int H = 32;               // hidden width
constexpr int kWaves = 2;           // waves per workgroup
constexpr int kTile = 32;           // samples per wave tile
constexpr int kPitch = 33;          // LDS row pitch of [hidden][sample] arrays

// Parameter vector layout.
constexpr int OFF_W1 = 0, OFF_B1 = 32, OFF_W2 = 64, OFF_B2 = 64 + 1024, OFF_W3 = 64 + 1024 + 32, OFF_B3 = 64 + 1024 + 64;
constexpr int NPARAM = OFF_B3 + 1;

// Logistic function with the hardware exp2 and reciprocal (about 1 ulp each);
// training and inference share it.
__device__ __forceinline__ float sigmoid(float z) { return __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(-1.44269504088896341f * z)); }

__device__ __forceinline__ void split8(const float *x, h8 &hi, h8 &lo)
{
#pragma unroll
	for (int i = 0; i < 8; ++i) {
		const _Float16 h = static_cast<_Float16>(x[i]);
		hi[i] = h;
		lo[i] = static_cast<_Float16>(x[i] - static_cast<float>(h));
	}
}

// D += A B with A, B given as FP32 operands (split into hi + lo here).
__device__ __forceinline__ f4 mma3(const float *a, const float *b, f4 acc)
{
	h8 ah, al, bh, bl;
	split8(a, ah, al);
	split8(b, bh, bl);
	acc = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, bh, acc, 0, 0, 0);
	acc = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, bl, acc, 0, 0, 0);
	acc = __builtin_amdgcn_mfma_f32_16x16x32_f16(al, bh, acc, 0, 0, 0);
	return acc;
}

struct WaveLds {
	float u[kTile], t[kTile], dv[kTile];
	float z1[H * kPitch], h1[H * kPitch], z2[H * kPitch], dz2[H * kPitch];
};

// Forward pass of one wave tile: U holds kTile inputs; leaves Z1, H1, Z2.
__device__ __forceinline__ void forward_tile(const float *p, WaveLds &w, unsigned lane)
{
	for (unsigned e = lane; e < H * kTile; e += 64) {
		const unsigned k = e / kTile, s = e % kTile;
		const float z = p[OFF_W1 + k] * w.u[s] + p[OFF_B1 + k];
		w.z1[k * kPitch + s] = z;
		w.h1[k * kPitch + s] = z * sigmoid(z);
	}
	__builtin_amdgcn_wave_barrier();
	__builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
	const unsigned row = lane % 16, kq = 8 * (lane / 16);
#pragma unroll
	for (unsigned c = 0; c < 2; ++c) {
		float a[8];
#pragma unroll
		for (int i = 0; i < 8; ++i)
			a[i] = p[OFF_W2 + (16 * c + row) * H + kq + i];
#pragma unroll
		for (unsigned g = 0; g < 2; ++g) {
			float b[8];
#pragma unroll
			for (int i = 0; i < 8; ++i)
				b[i] = w.h1[(kq + i) * kPitch + 16 * g + row];
			f4 acc = { 0.0f, 0.0f, 0.0f, 0.0f };
			acc = mma3(a, b, acc);
#pragma unroll
			for (int r = 0; r < 4; ++r) {
				const unsigned m = 16 * c + 4 * (lane / 16) + r;
				w.z2[m * kPitch + 16 * g + row] = acc[r] + p[OFF_B2 + m];
			}
		}
	}
	__builtin_amdgcn_wave_barrier();
	__builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
}

// Output of sample s after forward_tile.
__device__ __forceinline__ float output_of(const float *p, const WaveLds &w, unsigned s)
{
	float v = p[OFF_B3];
#pragma unroll 8
	for (int m = 0; m < H; ++m) {
		const float z = w.z2[m * kPitch + s];
		v += p[OFF_W3 + m] * (z * sigmoid(z));
	}
	return v;
}

// Network output of sample lane % kTile after forward_tile, in every lane of
// the wave (call from wave-uniform control flow): the two halves of the
// wave each sum 16 of the 32 hidden units of a sample, then exchange.
__device__ __forceinline__ float infer_output(const float *p, const WaveLds &w, unsigned lane)
{
	const unsigned s = lane % kTile, m0 = (lane / kTile) * (H / 2);
	float v = 0.0f;
#pragma unroll 8
	for (unsigned m = 0; m < H / 2; ++m) {
		const float z = w.z2[(m0 + m) * kPitch + s];
		v += p[OFF_W3 + m0 + m] * (z * sigmoid(z));
	}
	return p[OFF_B3] + (v + __shfl_xor(v, 32));
}

// One training pass over `tiles` wave tiles per wave: gradient of
// sum_s (v_s - t_s)^2 written as this workgroup's partial gradient, and the
// partial loss. Samples of workgroup b, wave w, tile i start at
// ((b * kWaves + w) * tiles + i) * kTile (mod n).
__global__ void __launch_bounds__(64 * kWaves) k_mlp_grad(const float *__restrict__ params, const float *__restrict__ xs,
                                                          const float *__restrict__ ts, unsigned n, unsigned offset, unsigned tiles,
                                                          float *__restrict__ partial)
{
	__shared__ float p[NPARAM];
	__shared__ float g[NPARAM + 1];
	__shared__ WaveLds wl[kWaves];
	const unsigned tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
	for (unsigned k = tid; k < NPARAM; k += 64 * kWaves) {
		p[k] = params[k];
		g[k] = 0.0f;
	}
	if (tid == 0)
		g[NPARAM] = 0.0f;
	__syncthreads();
	WaveLds &w = wl[wv];
	const unsigned row = lane % 16, kq = 8 * (lane / 16);
	f4 gw2[2][2] = {};
	float gw3[16] = {}, gb2[16] = {}, gw1[8] = {}, gb1[8] = {};
	float gb3 = 0.0f, loss = 0.0f;

	for (unsigned it = 0; it < tiles; ++it) {
		const unsigned base = ((blockIdx.x * kWaves + wv) * tiles + it) * kTile + offset;
		if (lane < kTile) {
			const unsigned s = (base + lane) % n;
			w.u[lane] = xs[s];
			w.t[lane] = ts[s];
		}
		__builtin_amdgcn_wave_barrier();
		__builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
		forward_tile(p, w, lane);
		if (lane < kTile) {
			const float v = output_of(p, w, lane);
			const float d = v - w.t[lane];
			loss += d * d;
			w.dv[lane] = 2.0f * d;
			gb3 += 2.0f * d;
		}
		__builtin_amdgcn_wave_barrier();
		__builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
		// dz2 = dv w3 silu'(z2); dw3 += dv h2; db2 += dz2
#pragma unroll
		for (unsigned q = 0; q < 16; ++q) {
			const unsigned e = lane + 64 * q, m = e / kTile, s = e % kTile;
			const float z = w.z2[m * kPitch + s];
			const float sg = sigmoid(z);
			const float dz = w.dv[s] * p[OFF_W3 + m] * (sg * (1.0f + z * (1.0f - sg)));
			w.dz2[m * kPitch + s] = dz;
			gw3[q] += w.dv[s] * (z * sg);
			gb2[q] += dz;
		}
		__builtin_amdgcn_wave_barrier();
		__builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
		// dW2 += dz2 h1^T (contraction over the 32 samples of the tile)
#pragma unroll
		for (unsigned c = 0; c < 2; ++c) {
			float a[8];
#pragma unroll
			for (int i = 0; i < 8; ++i)
				a[i] = w.dz2[(16 * c + row) * kPitch + kq + i];
#pragma unroll
			for (unsigned d = 0; d < 2; ++d) {
				float b[8];
#pragma unroll
				for (int i = 0; i < 8; ++i)
					b[i] = w.h1[(16 * d + row) * kPitch + kq + i];
				gw2[c][d] = mma3(a, b, gw2[c][d]);
			}
		}
		// dh1 = W2^T dz2, dz1 = dh1 silu'(z1); dw1 += dz1 u; db1 += dz1
#pragma unroll
		for (unsigned d = 0; d < 2; ++d) {
			float a[8];
#pragma unroll
			for (int i = 0; i < 8; ++i)
				a[i] = p[OFF_W2 + (kq + i) * H + 16 * d + row];
#pragma unroll
			for (unsigned gg = 0; gg < 2; ++gg) {
				float b[8];
#pragma unroll
				for (int i = 0; i < 8; ++i)
					b[i] = w.dz2[(kq + i) * kPitch + 16 * gg + row];
				f4 acc = { 0.0f, 0.0f, 0.0f, 0.0f };
				acc = mma3(a, b, acc);
#pragma unroll
				for (int r = 0; r < 4; ++r) {
					const unsigned j = 16 * d + 4 * (lane / 16) + r, s = 16 * gg + row;
					const float z = w.z1[j * kPitch + s];
					const float sg = sigmoid(z);
					const float dz = acc[r] * (sg * (1.0f + z * (1.0f - sg)));
					gw1[4 * d + r] += dz * w.u[s];
					gb1[4 * d + r] += dz;
				}
			}
		}
		__builtin_amdgcn_wave_barrier();
		__builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
	}

	// Workgroup reduction in LDS, then one partial vector per workgroup.
#pragma unroll
	for (unsigned q = 0; q < 16; ++q) {
		const unsigned m = (lane + 64 * q) / kTile;
		atomicAdd(&g[OFF_W3 + m], gw3[q]);
		atomicAdd(&g[OFF_B2 + m], gb2[q]);
	}
#pragma unroll
	for (unsigned d = 0; d < 2; ++d) {
#pragma unroll
		for (int r = 0; r < 4; ++r) {
			const unsigned j = 16 * d + 4 * (lane / 16) + r;
			atomicAdd(&g[OFF_W1 + j], gw1[4 * d + r]);
			atomicAdd(&g[OFF_B1 + j], gb1[4 * d + r]);
		}
	}
#pragma unroll
	for (unsigned c = 0; c < 2; ++c) {
#pragma unroll
		for (unsigned d = 0; d < 2; ++d) {
#pragma unroll
			for (int r = 0; r < 4; ++r) {
				const unsigned m = 16 * c + 4 * (lane / 16) + r, j = 16 * d + row;
				atomicAdd(&g[OFF_W2 + m * H + j], gw2[c][d][r]);
			}
		}
	}
	if (lane < kTile) {
		atomicAdd(&g[OFF_B3], gb3);
		atomicAdd(&g[NPARAM], loss);
	}
	__syncthreads();
	for (unsigned k = tid; k <= NPARAM; k += 64 * kWaves)
		partial[static_cast<size_t>(blockIdx.x) * (NPARAM + 1) + k] = g[k];
}

// Sums the partial gradients (scaled by 1/batch) and applies one Adam step
// (when lr > 0). grad_out (may be null) receives the reduced gradient and,
// in its last element, the mean loss.
__global__ void __launch_bounds__(256) k_mlp_adam(float *__restrict__ params, float *__restrict__ m1, float *__restrict__ m2,
                                                  const float *__restrict__ partial, unsigned blocks, float inv_batch, float lr,
                                                  float b1, float b2, float bc1, float bc2, float *__restrict__ grad_out)
{
	for (unsigned k = blockIdx.x * 256 + threadIdx.x; k <= static_cast<unsigned>(NPARAM); k += gridDim.x * 256) {
		float s = 0.0f;
		for (unsigned b = 0; b < blocks; ++b)
			s += partial[static_cast<size_t>(b) * (NPARAM + 1) + k];
		s *= inv_batch;
		if (grad_out)
			grad_out[k] = s;
		if (k == static_cast<unsigned>(NPARAM) || lr <= 0.0f)
			continue;
		const float a = b1 * m1[k] + (1.0f - b1) * s;
		const float v = b2 * m2[k] + (1.0f - b2) * s * s;
		m1[k] = a;
		m2[k] = v;
		params[k] -= lr * (a / bc1) / (sqrtf(v / bc2) + 1e-8f);
	}
}

// Applies the surrogate (or, outside its validated bins, the exact transfer)
// to n values in place: x -> post * f(x * pre).
__global__ void __launch_bounds__(64 * kWaves) k_surrogate_apply(float *__restrict__ x, unsigned n, learned::DeviceModel m)
{
	zm::lds_tables_init();
	__shared__ float p[NPARAM];
	__shared__ WaveLds wl[kWaves];
	const unsigned tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
	for (unsigned k = tid; k < NPARAM; k += 64 * kWaves)
		p[k] = m.params[k];
	__syncthreads();
	WaveLds &w = wl[wv];
	const unsigned waves = gridDim.x * kWaves;
	for (unsigned base = (blockIdx.x * kWaves + wv) * kTile; base < n; base += waves * kTile) {
		const unsigned s = base + lane;
		float xin = 0.0f, tin = 0.0f;
		bool use = false;
		if (lane < kTile && s < n) {
			xin = x[s];
			const float xs = xin * m.pre;
			tin = m.in_log ? __builtin_amdgcn_logf(xs) : xs;
			const float pos = (tin - m.lo) * m.inv_width;
			if (xs > 0.0f && pos >= 0.0f && pos < 1.0f) {
				const unsigned bin = min(static_cast<unsigned>(pos * m.nbins), m.nbins - 1);
				use = m.valid[bin] != 0;
			}
			w.u[lane] = use ? 2.0f * pos - 1.0f : 0.0f;
		}
		__builtin_amdgcn_wave_barrier();
		__builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
		// The network runs only for tiles with at least one validated value.
		float net = 0.0f;
		if (__builtin_amdgcn_ballot_w64(use) != 0) {
			forward_tile(p, w, lane);
			net = infer_output(p, w, lane);
		}
		if (lane < kTile && s < n) {
			float y;
			if (use) {
				const float v = net * m.out_scale + m.out_offset;
				y = m.post * (m.out_log ? __builtin_amdgcn_exp2f(v) : v);
			} else {
				y = m.post * zm::transfer(m.func, xin * m.pre);
			}
			x[s] = y;
		}
		__builtin_amdgcn_wave_barrier();
		__builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
	}
}

// The same over up to three pitched FLOAT planes of a batch (the surrogate
// form of a colourspace gamma operation): sample s of the flat index
// (plane, frame, row, column) space.
struct SurrogatePlanes {
	learned::PlaneView p[3];
	unsigned nplanes, width, height, frames;
	learned::DeviceModel m;
};

__global__ void __launch_bounds__(64 * kWaves) k_surrogate_planes(SurrogatePlanes a)
{
	zm::lds_tables_init();
	__shared__ float p[NPARAM];
	__shared__ WaveLds wl[kWaves];
	const learned::DeviceModel &m = a.m;
	const unsigned tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
	for (unsigned k = tid; k < NPARAM; k += 64 * kWaves)
		p[k] = m.params[k];
	__syncthreads();
	WaveLds &w = wl[wv];
	const uint64_t per_row = a.width, per_frame = per_row * a.height, per_plane = per_frame * a.frames;
	const uint64_t n = per_plane * a.nplanes;
	const uint64_t waves = static_cast<uint64_t>(gridDim.x) * kWaves;
	for (uint64_t base = (static_cast<uint64_t>(blockIdx.x) * kWaves + wv) * kTile; base < n; base += waves * kTile) {
		const uint64_t s = base + lane;
		float xin = 0.0f, tin = 0.0f;
		bool use = false;
		const float *src = nullptr;
		float *dst = nullptr;
		if (lane < kTile && s < n) {
			const unsigned q = static_cast<unsigned>(s / per_plane);
			const uint64_t r0 = s - q * per_plane;
			const unsigned f = static_cast<unsigned>(r0 / per_frame);
			const uint64_t r1 = r0 - f * per_frame;
			const unsigned row = static_cast<unsigned>(r1 / per_row), col = static_cast<unsigned>(r1 - row * per_row);
			const learned::PlaneView &v = a.p[q];
			src = reinterpret_cast<const float *>(v.src + f * v.src_fstride + row * v.src_stride) + col;
			dst = reinterpret_cast<float *>(v.dst + f * v.dst_fstride + row * v.dst_stride) + col;
			xin = *src;
			const float xs = xin * m.pre;
			tin = m.in_log ? __builtin_amdgcn_logf(xs) : xs;
			const float pos = (tin - m.lo) * m.inv_width;
			if (xs > 0.0f && pos >= 0.0f && pos < 1.0f) {
				const unsigned bin = min(static_cast<unsigned>(pos * m.nbins), m.nbins - 1);
				use = m.valid[bin] != 0;
			}
			w.u[lane] = use ? 2.0f * pos - 1.0f : 0.0f;
		}
		__builtin_amdgcn_wave_barrier();
		__builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
		float net = 0.0f;
		if (__builtin_amdgcn_ballot_w64(use) != 0) {
			forward_tile(p, w, lane);
			net = infer_output(p, w, lane);
		}
		if (dst) {
			float y;
			if (use) {
				const float v = net * m.out_scale + m.out_offset;
				y = m.post * (m.out_log ? __builtin_amdgcn_exp2f(v) : v);
			} else {
				y = m.post * zm::transfer(m.func, xin * m.pre);
			}
			*dst = y;
		}
		__builtin_amdgcn_wave_barrier();
		__builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
	}
}

// The exact transfer over n values in place (the method without the surrogate).
__global__ void __launch_bounds__(256) k_transfer_apply(float *__restrict__ x, unsigned n, int func, float pre, float post)
{
	zm::lds_tables_init();
	for (unsigned s = blockIdx.x * 256 + threadIdx.x; s < n; s += gridDim.x * 256)
		x[s] = post * zm::transfer(func, x[s] * pre);
}

// Raw network output (normalised units) at inputs u in [-1, 1].
__global__ void __launch_bounds__(64 * kWaves) k_mlp_eval(const float *__restrict__ params, const float *__restrict__ u, float *__restrict__ v,
                                                          unsigned n)
{
	__shared__ float p[NPARAM];
	__shared__ WaveLds wl[kWaves];
	const unsigned tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
	for (unsigned k = tid; k < NPARAM; k += 64 * kWaves)
		p[k] = params[k];
	__syncthreads();
	WaveLds &w = wl[wv];
	const unsigned waves = gridDim.x * kWaves;
	for (unsigned base = (blockIdx.x * kWaves + wv) * kTile; base < n; base += waves * kTile) {
		const unsigned s = base + lane;
		if (lane < kTile)
			w.u[lane] = s < n ? u[s] : 0.0f;
		__builtin_amdgcn_wave_barrier();
		__builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
		forward_tile(p, w, lane);
		const float out = infer_output(p, w, lane);
		if (lane < kTile && s < n)
			v[s] = out;
		__builtin_amdgcn_wave_barrier();
		__builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
	}
}

void check_hip(hipError_t e, const char *what)
{
	if (e != hipSuccess)
		fail(ZIMG_ERROR_UNKNOWN, std::string(what) + ": " + hipGetErrorString(e));
}

// ---- FP64 definitions of the transfer functions (the standards' formulas,
// gamma.cpp:63-338) used to generate training targets and to judge bins ----
constexpr double PQ_M1 = 0.1593017578125, PQ_M2 = 78.84375, PQ_C1 = 0.8359375, PQ_C2 = 18.8515625, PQ_C3 = 18.6875;
constexpr double HLG_A = 0.17883277, HLG_B = 0.28466892, HLG_C = 0.55991073;

double exact64(int func, double x)
{
	switch (func) {
	case zm::F_2084_EOTF: {
		if (!(x > 0)) return 0;
		const double xp = std::pow(x, 1.0 / PQ_M2);
		return std::pow(std::max(xp - PQ_C1, 0.0) / (PQ_C2 - PQ_C3 * xp), 1.0 / PQ_M1);
	}
	case zm::F_2084_INV_EOTF: {
		if (!(x > 0)) return 0;
		const double xp = std::pow(x, PQ_M1);
		return std::pow((PQ_C1 + PQ_C2 * xp) / (1.0 + PQ_C3 * xp), PQ_M2);
	}
	case zm::F_B67_OETF:
		x = std::max(x, 0.0);
		return x <= 1.0 / 12.0 ? std::sqrt(3.0 * x) : HLG_A * std::log(12.0 * x - HLG_B) + HLG_C;
	case zm::F_B67_INV_OETF:
		x = std::max(x, 0.0);
		return x <= 0.5 ? x * x / 3.0 : (std::exp((x - HLG_C) / HLG_A) + HLG_B) / 12.0;
	case zm::F_1886_EOTF: return x < 0 ? 0 : std::pow(x, 2.4);
	case zm::F_1886_INV_EOTF: return x < 0 ? 0 : std::pow(x, 1.0 / 2.4);
	default: return NAN;
	}
}

} // namespace

namespace learned {

int function_for(int transfer, int to_linear)
{
	switch (transfer) {
	case ZIMG_TRANSFER_ST2084: return to_linear ? zm::F_2084_EOTF : zm::F_2084_INV_EOTF;
	case ZIMG_TRANSFER_ARIB_B67: return to_linear ? zm::F_B67_INV_OETF : zm::F_B67_OETF;
	case ZIMG_TRANSFER_BT709:
	case ZIMG_TRANSFER_BT601:
	case ZIMG_TRANSFER_BT2020_10:
	case ZIMG_TRANSFER_BT2020_12: return to_linear ? zm::F_1886_EOTF : zm::F_1886_INV_EOTF;
	default: fail(ZIMG_ERROR_UNSUPPORTED_OPERATION, "no learned surrogate for this transfer function");
	}
}

double exact(int func, double x) { return exact64(func, x); }

int param_count() { return NPARAM; }

Surrogate::Surrogate(const Config &cfg, void *stream) : m_cfg(cfg)
{
	m_func = function_for(cfg.transfer, cfg.to_linear);
	// Input domain (in the transformed input space) and output transform.
	switch (m_func) {
	case zm::F_2084_EOTF: m_in_log = 0; m_lo = 0.05; m_hi = 1.0; m_out_log = 1; break;
	case zm::F_2084_INV_EOTF: m_in_log = 1; m_lo = -24.0; m_hi = 0.0; m_out_log = 0; break;
	case zm::F_B67_OETF: m_in_log = 1; m_lo = -24.0; m_hi = std::log2(12.0); m_out_log = 0; break;
	case zm::F_B67_INV_OETF: m_in_log = 0; m_lo = 0.02; m_hi = 1.0; m_out_log = 1; break;
	case zm::F_1886_EOTF: m_in_log = 0; m_lo = 0.0; m_hi = 1.0; m_out_log = 0; break;
	default: m_in_log = 0; m_lo = 0.0; m_hi = 1.0; m_out_log = 0; break;
	}
	hipStream_t s = static_cast<hipStream_t>(stream);

	// Training set: inputs uniform in the transformed domain, FP64 targets in
	// the transformed, normalised output space.
	const unsigned n = cfg.samples;
	std::vector<float> u(n), t(n);
	std::mt19937_64 rng(cfg.seed);
	std::uniform_real_distribution<double> uni(0.0, 1.0);
	std::vector<double> tv(n);
	double tmin = INFINITY, tmax = -INFINITY;
	for (unsigned k = 0; k < n; ++k) {
		const double pos = uni(rng);
		const double tin = m_lo + (m_hi - m_lo) * pos;
		const double x = m_in_log ? std::exp2(tin) : tin;
		double y = exact64(m_func, x);
		y = m_out_log ? std::log2(std::max(y, 1e-300)) : y;
		u[k] = static_cast<float>(2.0 * pos - 1.0);
		tv[k] = y;
		tmin = std::min(tmin, y);
		tmax = std::max(tmax, y);
	}
	m_out_offset = 0.5 * (tmin + tmax);
	m_out_scale = 0.5 * (tmax - tmin);
	if (!(m_out_scale > 0))
		m_out_scale = 1.0;
	for (unsigned k = 0; k < n; ++k)
		t[k] = static_cast<float>((tv[k] - m_out_offset) / m_out_scale);

	// Initial weights: uniform(-1/sqrt(fan_in), 1/sqrt(fan_in)).
	std::vector<float> p(NPARAM, 0.0f);
	auto init = [&](int off, int cnt, double fan) {
		std::uniform_real_distribution<double> d(-1.0 / std::sqrt(fan), 1.0 / std::sqrt(fan));
		for (int k = 0; k < cnt; ++k)
			p[off + k] = static_cast<float>(d(rng));
	};
	init(OFF_W1, 32, 1.0 / 9.0); // wider first layer spreads the SiLU knees over [-1, 1]
	init(OFF_B1, 32, 1.0 / 9.0);
	init(OFF_W2, 1024, 32);
	init(OFF_B2, 32, 32);
	init(OFF_W3, 32, 32);

	check_hip(hipMalloc(&m_params, NPARAM * sizeof(float)), "hipMalloc(surrogate)");
	check_hip(hipMemcpy(m_params, p.data(), NPARAM * sizeof(float), hipMemcpyHostToDevice), "hipMemcpy(surrogate)");
	float *xs = nullptr, *ts = nullptr, *m1 = nullptr, *m2 = nullptr, *partial = nullptr, *grad = nullptr;
	const unsigned blocks = 64, tiles = 4;
	const unsigned batch = blocks * kWaves * tiles * kTile;
	check_hip(hipMalloc(&xs, n * sizeof(float)), "hipMalloc");
	check_hip(hipMalloc(&ts, n * sizeof(float)), "hipMalloc");
	check_hip(hipMalloc(&m1, NPARAM * sizeof(float)), "hipMalloc");
	check_hip(hipMalloc(&m2, NPARAM * sizeof(float)), "hipMalloc");
	check_hip(hipMalloc(&partial, static_cast<size_t>(blocks) * (NPARAM + 1) * sizeof(float)), "hipMalloc");
	check_hip(hipMalloc(&grad, (NPARAM + 1) * sizeof(float)), "hipMalloc");
	check_hip(hipMemcpy(xs, u.data(), n * sizeof(float), hipMemcpyHostToDevice), "hipMemcpy");
	check_hip(hipMemcpy(ts, t.data(), n * sizeof(float), hipMemcpyHostToDevice), "hipMemcpy");
	check_hip(hipMemset(m1, 0, NPARAM * sizeof(float)), "hipMemset");
	check_hip(hipMemset(m2, 0, NPARAM * sizeof(float)), "hipMemset");

	const float b1 = 0.9f, b2 = 0.999f;
	double p1 = 1.0, p2 = 1.0;
	for (unsigned step = 0; step < cfg.steps; ++step) {
		p1 *= b1;
		p2 *= b2;
		const double frac = static_cast<double>(step) / std::max(1u, cfg.steps);
		const float lr = static_cast<float>(cfg.lr_min + 0.5 * (cfg.lr - cfg.lr_min) * (1.0 + std::cos(3.14159265358979323846 * frac)));
		const unsigned offset = static_cast<unsigned>((static_cast<uint64_t>(step) * batch) % n);
		hipLaunchKernelGGL(k_mlp_grad, dim3(blocks), dim3(64 * kWaves), 0, s, m_params, xs, ts, n, offset, tiles, partial);
		hipLaunchKernelGGL(k_mlp_adam, dim3(5), dim3(256), 0, s, m_params, m1, m2, partial, blocks, 1.0f / batch, lr, b1, b2,
		                   static_cast<float>(1.0 - p1), static_cast<float>(1.0 - p2), step + 1 == cfg.steps ? grad : nullptr);
	}
	check_hip(hipGetLastError(), "surrogate training launch");
	float last[NPARAM + 1];
	check_hip(hipMemcpyAsync(last, grad, sizeof(last), hipMemcpyDeviceToHost, s), "hipMemcpy");
	check_hip(hipStreamSynchronize(s), "hipStreamSynchronize");
	m_train_loss = last[NPARAM];
	(void)hipFree(xs);
	(void)hipFree(ts);
	(void)hipFree(m1);
	(void)hipFree(m2);
	(void)hipFree(partial);
	(void)hipFree(grad);

	validate(stream);
}

Surrogate::~Surrogate()
{
	if (m_params)
		(void)hipFree(m_params);
	if (m_valid)
		(void)hipFree(m_valid);
}

// Bins of the input domain where the device inference, at 64 points per bin
// (both ends included), stays within the tolerance of the FP64 function.
void Surrogate::validate(void *stream)
{
	hipStream_t s = static_cast<hipStream_t>(stream);
	const unsigned nb = m_cfg.bins, per = 64;
	const unsigned n = nb * per;
	std::vector<float> xv(n);
	for (unsigned b = 0; b < nb; ++b) {
		for (unsigned k = 0; k < per; ++k) {
			const double pos = (b + static_cast<double>(k) / (per - 1) * 0.999999) / nb;
			const double tin = m_lo + (m_hi - m_lo) * pos;
			xv[static_cast<size_t>(b) * per + k] = static_cast<float>(m_in_log ? std::exp2(tin) : tin);
		}
	}
	std::vector<uint8_t> ones(nb, 1);
	check_hip(hipMalloc(&m_valid, nb), "hipMalloc");
	check_hip(hipMemcpy(m_valid, ones.data(), nb, hipMemcpyHostToDevice), "hipMemcpy");
	float *dx = nullptr;
	check_hip(hipMalloc(&dx, n * sizeof(float)), "hipMalloc");
	check_hip(hipMemcpy(dx, xv.data(), n * sizeof(float), hipMemcpyHostToDevice), "hipMemcpy");
	// Unit pre/post scaling for validation: the tolerance is judged on f itself.
	DeviceModel m = device_model(1.0f, 1.0f);
	hipLaunchKernelGGL(k_surrogate_apply, dim3(256), dim3(64 * kWaves), 0, s, dx, n, m);
	check_hip(hipGetLastError(), "surrogate validation launch");
	std::vector<float> yv(n);
	check_hip(hipMemcpyAsync(yv.data(), dx, n * sizeof(float), hipMemcpyDeviceToHost, s), "hipMemcpy");
	check_hip(hipStreamSynchronize(s), "hipStreamSynchronize");
	(void)hipFree(dx);
	std::vector<uint8_t> valid(nb, 0);
	m_coverage = 0;
	m_max_err = 0;
	for (unsigned b = 0; b < nb; ++b) {
		double worst = 0;
		for (unsigned k = 0; k < per; ++k) {
			const size_t i = static_cast<size_t>(b) * per + k;
			const double want = exact64(m_func, static_cast<double>(xv[i]));
			double err = std::fabs(static_cast<double>(yv[i]) - want);
			if (m_cfg.relative)
				err /= std::max(std::fabs(want), 1e-300);
			if (std::isnan(err))
				err = INFINITY;
			worst = std::max(worst, err);
		}
		if (worst <= m_cfg.tolerance) {
			valid[b] = 1;
			++m_coverage;
			m_max_err = std::max(m_max_err, worst);
		}
	}
	check_hip(hipMemcpy(m_valid, valid.data(), nb, hipMemcpyHostToDevice), "hipMemcpy");
}

DeviceModel Surrogate::device_model(float pre, float post) const
{
	DeviceModel m{};
	m.params = m_params;
	m.valid = m_valid;
	m.nbins = m_cfg.bins;
	m.in_log = m_in_log;
	m.out_log = m_out_log;
	m.lo = static_cast<float>(m_lo);
	m.inv_width = static_cast<float>(1.0 / (m_hi - m_lo));
	m.out_scale = static_cast<float>(m_out_scale);
	m.out_offset = static_cast<float>(m_out_offset);
	m.func = m_func;
	m.pre = pre;
	m.post = post;
	return m;
}

void Surrogate::apply(float *x, size_t n, float pre, float post, void *stream) const
{
	if (n > 0xFFFFFFFFu)
		fail(ZIMG_ERROR_ILLEGAL_ARGUMENT, "too many values");
	const DeviceModel m = device_model(pre, post);
	hipLaunchKernelGGL(k_surrogate_apply, dim3(2048), dim3(64 * kWaves), 0, static_cast<hipStream_t>(stream), x, static_cast<unsigned>(n), m);
	check_hip(hipGetLastError(), "surrogate launch");
}

void apply_exact(int func, float *x, size_t n, float pre, float post, void *stream)
{
	if (n > 0xFFFFFFFFu)
		fail(ZIMG_ERROR_ILLEGAL_ARGUMENT, "too many values");
	hipLaunchKernelGGL(k_transfer_apply, dim3(4096), dim3(256), 0, static_cast<hipStream_t>(stream), x, static_cast<unsigned>(n), func, pre,
	                   post);
	check_hip(hipGetLastError(), "transfer launch");
}

void apply_planes(const DeviceModel &m, const PlaneView *planes, unsigned nplanes, unsigned width, unsigned height, unsigned frames,
                  void *stream)
{
	if (nplanes > 3)
		fail(ZIMG_ERROR_ILLEGAL_ARGUMENT, "at most three planes");
	SurrogatePlanes a{};
	for (unsigned q = 0; q < nplanes; ++q)
		a.p[q] = planes[q];
	a.nplanes = nplanes;
	a.width = width;
	a.height = height;
	a.frames = frames;
	a.m = m;
	hipLaunchKernelGGL(k_surrogate_planes, dim3(4096), dim3(64 * kWaves), 0, static_cast<hipStream_t>(stream), a);
	check_hip(hipGetLastError(), "surrogate launch");
}

void Surrogate::get_params(float *out) const
{
	check_hip(hipMemcpy(out, m_params, NPARAM * sizeof(float), hipMemcpyDeviceToHost), "hipMemcpy");
}

void gradients(const float *params, const float *u, const float *t, unsigned n, float *grad, void *stream)
{
	if (n == 0 || n % (kWaves * kTile))
		fail(ZIMG_ERROR_ILLEGAL_ARGUMENT, "sample count must be a positive multiple of 64");
	hipStream_t s = static_cast<hipStream_t>(stream);
	const unsigned blocks = n / (kWaves * kTile);
	float *dp = nullptr, *du = nullptr, *dt = nullptr, *partial = nullptr, *dg = nullptr;
	check_hip(hipMalloc(&dp, NPARAM * sizeof(float)), "hipMalloc");
	check_hip(hipMalloc(&du, n * sizeof(float)), "hipMalloc");
	check_hip(hipMalloc(&dt, n * sizeof(float)), "hipMalloc");
	check_hip(hipMalloc(&partial, static_cast<size_t>(blocks) * (NPARAM + 1) * sizeof(float)), "hipMalloc");
	check_hip(hipMalloc(&dg, (NPARAM + 1) * sizeof(float)), "hipMalloc");
	check_hip(hipMemcpy(dp, params, NPARAM * sizeof(float), hipMemcpyHostToDevice), "hipMemcpy");
	check_hip(hipMemcpy(du, u, n * sizeof(float), hipMemcpyHostToDevice), "hipMemcpy");
	check_hip(hipMemcpy(dt, t, n * sizeof(float), hipMemcpyHostToDevice), "hipMemcpy");
	hipLaunchKernelGGL(k_mlp_grad, dim3(blocks), dim3(64 * kWaves), 0, s, dp, du, dt, n, 0u, 1u, partial);
	hipLaunchKernelGGL(k_mlp_adam, dim3(5), dim3(256), 0, s, dp, nullptr, nullptr, partial, blocks, 1.0f / n, 0.0f, 0.0f, 0.0f, 1.0f, 1.0f, dg);
	check_hip(hipGetLastError(), "gradient launch");
	check_hip(hipMemcpyAsync(grad, dg, (NPARAM + 1) * sizeof(float), hipMemcpyDeviceToHost, s), "hipMemcpy");
	check_hip(hipStreamSynchronize(s), "hipStreamSynchronize");
	(void)hipFree(dp);
	(void)hipFree(du);
	(void)hipFree(dt);
	(void)hipFree(partial);
	(void)hipFree(dg);
}

void eval(const float *params, const float *u, float *v, unsigned n, void *stream)
{
	hipStream_t s = static_cast<hipStream_t>(stream);
	float *dp = nullptr, *du = nullptr, *dv = nullptr;
	check_hip(hipMalloc(&dp, NPARAM * sizeof(float)), "hipMalloc");
	check_hip(hipMalloc(&du, n * sizeof(float)), "hipMalloc");
	check_hip(hipMalloc(&dv, n * sizeof(float)), "hipMalloc");
	check_hip(hipMemcpy(dp, params, NPARAM * sizeof(float), hipMemcpyHostToDevice), "hipMemcpy");
	check_hip(hipMemcpy(du, u, n * sizeof(float), hipMemcpyHostToDevice), "hipMemcpy");
	hipLaunchKernelGGL(k_mlp_eval, dim3(64), dim3(64 * kWaves), 0, s, dp, du, dv, n);
	check_hip(hipGetLastError(), "eval launch");
	check_hip(hipMemcpyAsync(v, dv, n * sizeof(float), hipMemcpyDeviceToHost, s), "hipMemcpy");
	check_hip(hipStreamSynchronize(s), "hipStreamSynchronize");
	(void)hipFree(dp);
	(void)hipFree(du);
	(void)hipFree(dv);
}

} // namespace learned
} // namespace zamd
