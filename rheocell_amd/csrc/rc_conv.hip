// rc_conv.hip -- BF16-MFMA 3x3 convolutions (circular padding) for the
// learned CG warm start: forward, backward-data, backward-weights, Adam.
//
// Network: 6 conv layers, 32 channels, 3x3, circular padding, ReLU between
// layers.  Every tensor is NHWC with 32 bf16 channels (the 3 input and 2
// output channels are zero-padded to 32), so one kernel shape serves every
// layer: an implicit GEMM in which one 64-lane wave computes 16 pixels x 32
// output channels with v_mfma_f32_16x16x32_bf16 (K = 32 input channels of one
// tap per instruction, 9 taps, 2 output-channel tiles), accumulating in FP32.
//
// MFMA 16x16x32 bf16 operand layout (lane l, element j):
//   A[i][k]: i = l % 16, k = 8 * (l / 16) + j        (8 bf16 per lane)
//   B[k][n]: n = l % 16, k = 8 * (l / 16) + j        (8 bf16 per lane)
//   D[i][n]: n = l % 16, i = 4 * (l / 16) + r        (4 fp32 per lane)
#include <hip/hip_runtime.h>

#include "rc_conv.h"

namespace rc {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));

__device__ inline bf16x8 load8(const unsigned short *p) {
	const u16x8 v = *reinterpret_cast<const u16x8 *>(p);
	return __builtin_bit_cast(bf16x8, v);
}

// out[n, y, x, co] = act(bias[co] + sum_{tap, ci} in[n, y+dy, x+dx, ci] * w[tap][co][ci]),
// (dy, dx) = (tap / 3 - 1, tap % 3 - 1).  Backprop through a layer uses the
// same kernel with the tap-mirrored, transposed weights wt[8 - tap][ci][co].
// mode: 0 store bf16 ReLU(z); 1 store bf16 z; 2 store bf16 z masked by (mask > 0) (backprop through ReLU);
// out may be null (no bf16 store); fout (optional): fp32 copy of channels 0..1 of z.
//
// A workgroup (CONV_ROWS waves) computes a CONV_ROWS x (16 CONV_TILES) pixel
// block: wave w takes image row y0 + w and its CONV_TILES 16-pixel tiles, with
// the layer's weights held in registers (9 taps x 2 output-channel halves = 18
// A operands).  The weights are staged in LDS once per workgroup and the
// block's input rows y0-1 .. y0+CONV_ROWS and columns x0-1 .. x0+16 CONV_TILES
// (circular) once as well, so every input pixel leaves HBM/L2 about
// (CONV_ROWS+2)/CONV_ROWS times instead of 9 times.
// LDS rows are 64 bytes (32 bf16) with the 16-byte chunk index XOR-swizzled by
// bits 2..3 of the row index: the 16 lanes of a 128-bit read of 16
// consecutive rows then hit distinct banks.  Weights are the A operand and
// pixels the B operand, so D[co][px] leaves each lane 4 consecutive channels
// of one pixel: bf16 results go out as 8-byte stores.
constexpr int CONV_ROWS = 8;                       // image rows per workgroup, one per wave
constexpr int CONV_THREADS = 64 * CONV_ROWS;
constexpr int CONV_TILES = 4;
constexpr int CONV_BW = 16 * CONV_TILES;           // block width in pixels
constexpr int CONV_SW = CONV_BW + 2;               // staged columns
constexpr int CONV_SH = CONV_ROWS + 2;             // staged rows
constexpr int CONV_NIN = CONV_SH * CONV_SW * 4;     // staged 16-byte input chunks
constexpr int CONV_LIN = (CONV_NIN + CONV_THREADS - 1) / CONV_THREADS;   // per thread
constexpr int CONV_NW = 9 * 32 * 4;                // 16-byte weight chunks of a layer
constexpr int CONV_LW = (CONV_NW + CONV_THREADS - 1) / CONV_THREADS;

// two floats to packed bf16, rounded to nearest even by v_cvt_pk_bf16_f32
__device__ inline unsigned pack2bf(float a, float b) {
	return (unsigned) __builtin_bit_cast(unsigned short, (__bf16) a) |
			((unsigned) __builtin_bit_cast(unsigned short, (__bf16) b) << 16);
}
// element offset of 16-byte chunk c (0..3) of LDS row q
__device__ inline int swz(int q, int c) { return q * 32 + ((c ^ ((q >> 2) & 3)) << 3); }

// MODE and OUT (bf16 store wanted) are compile-time; fout may be null.
// run: if set and *run == 0, the launch does nothing (a learned CG stage
// after the solve has converged)
template <int MODE, bool OUT>
__global__ __launch_bounds__(CONV_THREADS) void k_conv3x3(const unsigned short *__restrict__ in,
		const unsigned short *__restrict__ w, const float *__restrict__ bias, unsigned short *__restrict__ out,
		const unsigned short *__restrict__ mask, float *__restrict__ fout, int H, int W, const int *run) {
	if(run && *run == 0) return;
	__shared__ __attribute__((aligned(16))) unsigned short s_in[CONV_SH * CONV_SW * 32];
	__shared__ __attribute__((aligned(16))) unsigned short s_w[9 * 32 * 32];   // rows (tap, co), swizzled like s_in
	const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
	const int bx = (W + CONV_BW - 1) / CONV_BW;
	const int x0 = (blockIdx.x % bx) * CONV_BW, y0 = (blockIdx.x / bx) * CONV_ROWS;
	const int n = blockIdx.y;
	const size_t img = (size_t) n * H * W;
	// stage: CONV_SH rows x CONV_SW columns x 4 chunks of 8 channels, and the
	// layer's weights (once per workgroup; every wave reads them from LDS).
	// Every load of a thread is issued before the first LDS write, so the
	// staging costs one memory round trip, not one per chunk.
	uint4 vin[CONV_LIN], vw[CONV_LW];
#pragma unroll
	for(int j = 0; j < CONV_LIN; ++j) {
		// past the end (last round only): the last chunk again
		const int e = min((int) threadIdx.x + j * CONV_THREADS, CONV_NIN - 1);
		const int ch = e & 3, q = e >> 2, col = q % CONV_SW, r = q / CONV_SW;
		// circular (operands lie in [-1, 2W) and [-1, 2H) as W >= 32 >= 16 CONV_TILES / 2 and
		// H >= 3); columns past W of a partial block wrap too (staged, never read)
		int xs = x0 + col - 1, ys = y0 + r - 1;
		xs += xs < 0 ? W : 0;
		xs -= xs >= W ? W : 0;
		xs -= xs >= W ? W : 0;
		ys += ys < 0 ? H : 0;
		ys -= ys >= H ? H : 0;
		ys -= ys >= H ? H : 0;
		vin[j] = *reinterpret_cast<const uint4 *>(in + (img + (size_t) ys * W + xs) * 32 + ch * 8);
	}
#pragma unroll
	for(int j = 0; j < CONV_LW; ++j) {
		const int e = min((int) threadIdx.x + j * CONV_THREADS, CONV_NW - 1);
		vw[j] = *reinterpret_cast<const uint4 *>(w + (size_t) e * 8);
	}
	const int i = lane & 15, kb = lane >> 4;
	float bz0[4], bz1[4];
#pragma unroll
	for(int r = 0; r < 4; ++r) {
		bz0[r] = bias ? bias[4 * kb + r] : 0.f;
		bz1[r] = bias ? bias[16 + 4 * kb + r] : 0.f;
	}
#pragma unroll
	for(int j = 0; j < CONV_LIN; ++j) {
		// unconditional: a thread past the end writes the last chunk's own value
		// again (a guarded write would let the compiler sink its load behind it)
		const int e = min((int) threadIdx.x + j * CONV_THREADS, CONV_NIN - 1);
		*reinterpret_cast<uint4 *>(&s_in[swz(e >> 2, e & 3)]) = vin[j];
	}
#pragma unroll
	for(int j = 0; j < CONV_LW; ++j) {
		const int e = min((int) threadIdx.x + j * CONV_THREADS, CONV_NW - 1);
		*reinterpret_cast<uint4 *>(&s_w[swz(e >> 2, e & 3)]) = vw[j];
	}
	__syncthreads();
	bf16x8 wa[9], wb[9];
#pragma unroll
	for(int tap = 0; tap < 9; ++tap) {
		wa[tap] = __builtin_bit_cast(bf16x8, *reinterpret_cast<const u16x8 *>(&s_w[swz(tap * 32 + i, kb)]));
		wb[tap] = __builtin_bit_cast(bf16x8, *reinterpret_cast<const u16x8 *>(&s_w[swz(tap * 32 + 16 + i, kb)]));
	}
	const int y = y0 + wave;
	if(y >= H) return;
	// LDS element offset of each tap's operand for tile 0: staged row wave + dy,
	// column i + dx.  Tile t adds t * 16 rows of 32 elements: the swizzle depends
	// on bits 2..3 of the row index, which a step of 16 leaves unchanged.
	int tap_off[9];
#pragma unroll
	for(int tap = 0; tap < 9; ++tap) tap_off[tap] = swz((wave + tap / 3) * CONV_SW + i + tap % 3, kb);
#pragma unroll
	for(int t = 0; t < CONV_TILES; ++t) {
		const int px = x0 + t * 16 + i;
		if(x0 + t * 16 >= W) break;
		f32x4 acc0 = { 0.f, 0.f, 0.f, 0.f }, acc1 = { 0.f, 0.f, 0.f, 0.f };
		bf16x8 a[9];
#pragma unroll
		for(int tap = 0; tap < 9; ++tap) {
			const u16x8 v = *reinterpret_cast<const u16x8 *>(&s_in[tap_off[tap] + t * 16 * 32]);
			a[tap] = __builtin_bit_cast(bf16x8, v);
		}
#pragma unroll
		for(int tap = 0; tap < 9; ++tap) {
			acc0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wa[tap], a[tap], acc0, 0, 0, 0);
			acc1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wb[tap], a[tap], acc1, 0, 0, 0);
		}
		// D[co][px]: lane holds channels 4 kb + r (acc0) and 16 + 4 kb + r (acc1) of pixel px
		const size_t pix = img + (size_t) y * W + px;
		float z0[4], z1[4];
#pragma unroll
		for(int r = 0; r < 4; ++r) {
			z0[r] = acc0[r] + bz0[r];
			z1[r] = acc1[r] + bz1[r];
		}
		if(fout && kb == 0) *reinterpret_cast<float2 *>(fout + pix * 2) = make_float2(z0[0], z0[1]);
		if(!OUT) continue;
		if(MODE == 0) {
#pragma unroll
			for(int r = 0; r < 4; ++r) {
				z0[r] = fmaxf(z0[r], 0.f);
				z1[r] = fmaxf(z1[r], 0.f);
			}
		}
		else if(MODE == 2) {
			const uint2 m0 = *reinterpret_cast<const uint2 *>(mask + pix * 32 + 4 * kb);
			const uint2 m1 = *reinterpret_cast<const uint2 *>(mask + pix * 32 + 16 + 4 * kb);
			const unsigned mm0[4] = { m0.x & 0xFFFFu, m0.x >> 16, m0.y & 0xFFFFu, m0.y >> 16 };
			const unsigned mm1[4] = { m1.x & 0xFFFFu, m1.x >> 16, m1.y & 0xFFFFu, m1.y >> 16 };
#pragma unroll
			for(int r = 0; r < 4; ++r) {
				if(!(bf2f((unsigned short) mm0[r]) > 0.f)) z0[r] = 0.f;
				if(!(bf2f((unsigned short) mm1[r]) > 0.f)) z1[r] = 0.f;
			}
		}
		*reinterpret_cast<uint2 *>(out + pix * 32 + 4 * kb) = make_uint2(pack2bf(z0[0], z0[1]), pack2bf(z0[2], z0[3]));
		*reinterpret_cast<uint2 *>(out + pix * 32 + 16 + 4 * kb) = make_uint2(pack2bf(z1[0], z1[1]), pack2bf(z1[2], z1[3]));
	}
}

hipError_t launch_conv3x3(const unsigned short *in, const unsigned short *w, const float *bias, unsigned short *out,
		const unsigned short *mask, float *fout, int N, int H, int W, int mode, hipStream_t s, const int *run) {
	if(W % 16 || W < 32 || H < 3) return hipErrorInvalidValue;   // the staging wraps assume these
	const int blocks = ((W + CONV_BW - 1) / CONV_BW) * ((H + CONV_ROWS - 1) / CONV_ROWS);
	const dim3 g(blocks, N);
	const dim3 b(CONV_THREADS);
	if(!out) hipLaunchKernelGGL((k_conv3x3<1, false>), g, b, 0, s, in, w, bias, out, mask, fout, H, W, run);
	else if(mode == 0) hipLaunchKernelGGL((k_conv3x3<0, true>), g, b, 0, s, in, w, bias, out, mask, fout, H, W, run);
	else if(mode == 2) hipLaunchKernelGGL((k_conv3x3<2, true>), g, b, 0, s, in, w, bias, out, mask, fout, H, W, run);
	else hipLaunchKernelGGL((k_conv3x3<1, true>), g, b, 0, s, in, w, bias, out, mask, fout, H, W, run);
	return hipGetLastError();
}

// Weight gradient partials: for a chunk of ROWS image rows of one sample,
// dW[tap][co][ci] += sum_p g[p][co] * in[p + tap][ci], db[co] += sum_p g[p][co].
// Staged through LDS: rows y-1..y+1 of the input (with circular x halo) and
// row y of g, 32 pixels at a time; wave w computes the (co, ci) tile
// (w / 2, w % 2) for all 9 taps.
constexpr int WG_ROWS = 8;
__global__ __launch_bounds__(256) void k_conv_wgrad(const unsigned short *__restrict__ in, const unsigned short *__restrict__ g,
		float *__restrict__ part, int H, int W) {
	__shared__ unsigned short s_in[3][34][32];
	__shared__ unsigned short s_g[32][32];
	const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
	const int mt = wave >> 1, nt = wave & 1;
	const int i = lane & 15, kb = lane >> 4;
	const int chunks_per_img = (H + WG_ROWS - 1) / WG_ROWS;
	const int n = blockIdx.x / chunks_per_img, y0 = (blockIdx.x % chunks_per_img) * WG_ROWS;
	const size_t img = (size_t) n * H * W;
	f32x4 acc[9];
#pragma unroll
	for(int t = 0; t < 9; ++t) acc[t] = f32x4{ 0.f, 0.f, 0.f, 0.f };
	float dbias = 0.f;
	for(int y = y0; y < min(y0 + WG_ROWS, H); ++y) {
		for(int xb = 0; xb < W; xb += 32) {
			__syncthreads();
			// stage input rows y-1..y+1, pixels xb-1 .. xb+32 (circular), and g row y
			for(int e = threadIdx.x; e < 3 * 34 * 4; e += 256) {
				const int part8 = e & 3, px = (e >> 2) % 34, ry = (e >> 2) / 34;
				int xs = xb + px - 1;
				xs = xs < 0 ? xs + W : (xs >= W ? xs - W : xs);
				int ys = y + ry - 1;
				ys = ys < 0 ? ys + H : (ys >= H ? ys - H : ys);
				*reinterpret_cast<u16x8 *>(&s_in[ry][px][part8 * 8]) =
						*reinterpret_cast<const u16x8 *>(in + (img + (size_t) ys * W + xs) * 32 + part8 * 8);
			}
			for(int e = threadIdx.x; e < 32 * 4; e += 256) {
				const int part8 = e & 3, px = e >> 2;
				u16x8 v = u16x8{ 0, 0, 0, 0, 0, 0, 0, 0 };
				if(xb + px < W) v = *reinterpret_cast<const u16x8 *>(g + (img + (size_t) y * W + xb + px) * 32 + part8 * 8);
				*reinterpret_cast<u16x8 *>(&s_g[px][part8 * 8]) = v;
			}
			__syncthreads();
			// A[co][k = pixel]: co = 16*mt + i, pixels 8*kb + j ; B[k = pixel][ci]: ci = 16*nt + i
			u16x8 av;
#pragma unroll
			for(int j = 0; j < 8; ++j) av[j] = s_g[8 * kb + j][16 * mt + i];
			const bf16x8 a = __builtin_bit_cast(bf16x8, av);
			if(nt == 0 && kb == 0) {
				// bias gradient of channel 16*mt + i over the 32 pixels
				for(int j = 0; j < 32; ++j) dbias += bf2f(s_g[j][16 * mt + i]);
			}
#pragma unroll
			for(int t = 0; t < 9; ++t) {
				const int dy = t / 3 - 1, dx = t % 3 - 1;
				u16x8 bv;
#pragma unroll
				for(int j = 0; j < 8; ++j) bv[j] = s_in[dy + 1][8 * kb + j + dx + 1][16 * nt + i];
				acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, __builtin_bit_cast(bf16x8, bv), acc[t], 0, 0, 0);
			}
		}
	}
	// part[blk][tap][co][ci] then part_bias[blk][co] at offset 9*32*32
	float *P = part + (size_t) blockIdx.x * (9 * 1024 + 32);
#pragma unroll
	for(int t = 0; t < 9; ++t)
#pragma unroll
		for(int r = 0; r < 4; ++r) P[t * 1024 + (16 * mt + 4 * kb + r) * 32 + 16 * nt + i] = acc[t][r];
	if(nt == 0 && kb == 0) P[9 * 1024 + 16 * mt + i] = dbias;
}

hipError_t launch_conv_wgrad(const unsigned short *in, const unsigned short *g, float *part, int N, int H, int W, hipStream_t s) {
	const int blocks = N * ((H + WG_ROWS - 1) / WG_ROWS);
	hipLaunchKernelGGL(k_conv_wgrad, dim3(blocks), dim3(256), 0, s, in, g, part, H, W);
	return hipGetLastError();
}

int conv_wgrad_blocks(int N, int H) { return N * ((H + WG_ROWS - 1) / WG_ROWS); }

// sum the partials in a fixed order: grad[e] = sum_b part[b][e]
__global__ void k_sum_parts(const float *__restrict__ part, float *__restrict__ grad, int nblk, int len) {
	const int e = blockIdx.x * 256 + threadIdx.x;
	if(e >= len) return;
	float s = 0.f;
	for(int b = 0; b < nblk; ++b) s += part[(size_t) b * len + e];
	grad[e] = s;
}

hipError_t launch_sum_parts(const float *part, float *grad, int nblk, int len, hipStream_t s) {
	hipLaunchKernelGGL(k_sum_parts, dim3((len + 255) / 256), dim3(256), 0, s, part, grad, nblk, len);
	return hipGetLastError();
}

// Adam on the fp32 master copy; refreshes the bf16 weights (w[tap][co][ci]) and
// the mirrored/transposed copy for backprop (wt[tap'][ci][co], tap' = 8 - tap).
// mask[e] == 0 keeps padded entries at zero.
__global__ void k_adam(float *__restrict__ p, float *__restrict__ m, float *__restrict__ v, const float *__restrict__ gr,
		const float *__restrict__ mask, int len, float lr, float b1, float b2, float eps, float c1, float c2, float scale) {
	const int e = blockIdx.x * 256 + threadIdx.x;
	if(e >= len) return;
	const float g = gr[e] * scale * mask[e];
	const float mm = b1 * m[e] + (1.f - b1) * g;
	const float vv = b2 * v[e] + (1.f - b2) * g * g;
	m[e] = mm;
	v[e] = vv;
	p[e] = (p[e] - lr * (mm / c1) / (sqrtf(vv / c2) + eps)) * mask[e];
}

hipError_t launch_adam(float *p, float *m, float *v, const float *g, const float *mask, int len, float lr, float b1,
		float b2, float eps, int t, float scale, hipStream_t s) {
	const float c1 = 1.f - powf(b1, (float) t), c2 = 1.f - powf(b2, (float) t);
	hipLaunchKernelGGL(k_adam, dim3((len + 255) / 256), dim3(256), 0, s, p, m, v, g, mask, len, lr, b1, b2, eps, c1, c2, scale);
	return hipGetLastError();
}

__global__ void k_weights_to_bf16(const float *__restrict__ wf, unsigned short *__restrict__ w, unsigned short *__restrict__ wt, int layers) {
	const int e = blockIdx.x * 256 + threadIdx.x;
	if(e >= layers * 9 * 1024) return;
	const int l = e / (9 * 1024), r = e % (9 * 1024), tap = r / 1024, co = (r % 1024) / 32, ci = r % 32;
	const unsigned short b = f2bf(wf[e]);
	w[e] = b;
	wt[(size_t) l * 9 * 1024 + (8 - tap) * 1024 + ci * 32 + co] = b;
}

hipError_t launch_weights_to_bf16(const float *wf, unsigned short *w, unsigned short *wt, int layers, hipStream_t s) {
	hipLaunchKernelGGL(k_weights_to_bf16, dim3((layers * 9 * 1024 + 255) / 256), dim3(256), 0, s, wf, w, wt, layers);
	return hipGetLastError();
}

// x: fp32 [n][p][3] features -> bf16 NHWC32 (channels >= 3 zero)
__global__ void k_pack_features(const float *__restrict__ x, unsigned short *__restrict__ act, long long npix) {
	const long long p = (long long) blockIdx.x * 256 + threadIdx.x;
	if(p >= npix) return;
	u16x8 v0 = u16x8{ 0, 0, 0, 0, 0, 0, 0, 0 };
	v0[0] = f2bf(x[p * 3 + 0]);
	v0[1] = f2bf(x[p * 3 + 1]);
	v0[2] = f2bf(x[p * 3 + 2]);
	u16x8 *o = reinterpret_cast<u16x8 *>(act + p * 32);
	o[0] = v0;
	o[1] = o[2] = o[3] = u16x8{ 0, 0, 0, 0, 0, 0, 0, 0 };
}

hipError_t launch_pack_features(const float *x, unsigned short *act, long long npix, hipStream_t s) {
	hipLaunchKernelGGL(k_pack_features, dim3((unsigned) ((npix + 255) / 256)), dim3(256), 0, s, x, act, npix);
	return hipGetLastError();
}

// MSE loss on channels 0..1: partial sums per block, and the bf16 gradient
// g = 2 (z - y) / count in NHWC32 (other channels zero)
__global__ void k_mse(const float *__restrict__ z, const float *__restrict__ y, unsigned short *__restrict__ g,
		double *__restrict__ part, long long npix, float inv_count) {
	__shared__ double s[256];
	const long long p = (long long) blockIdx.x * 256 + threadIdx.x;
	double l = 0.;
	if(p < npix) {
		const float e0 = z[p * 2] - y[p * 2], e1 = z[p * 2 + 1] - y[p * 2 + 1];
		l = (double) e0 * e0 + (double) e1 * e1;
		u16x8 v0 = u16x8{ 0, 0, 0, 0, 0, 0, 0, 0 };
		v0[0] = f2bf(2.f * e0 * inv_count);
		v0[1] = f2bf(2.f * e1 * inv_count);
		u16x8 *o = reinterpret_cast<u16x8 *>(g + p * 32);
		o[0] = v0;
		o[1] = o[2] = o[3] = u16x8{ 0, 0, 0, 0, 0, 0, 0, 0 };
	}
	s[threadIdx.x] = l;
	__syncthreads();
	for(int k = 128; k > 0; k >>= 1) {
		if(threadIdx.x < k) s[threadIdx.x] += s[threadIdx.x + k];
		__syncthreads();
	}
	if(threadIdx.x == 0) part[blockIdx.x] = s[0];
}

hipError_t launch_mse(const float *z, const float *y, unsigned short *g, double *part, long long npix, float inv_count,
		hipStream_t s) {
	hipLaunchKernelGGL(k_mse, dim3((unsigned) ((npix + 255) / 256)), dim3(256), 0, s, z, y, g, part, npix, inv_count);
	return hipGetLastError();
}

}  // namespace rc
