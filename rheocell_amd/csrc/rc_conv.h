// rc_conv.h -- launch interface of the BF16-MFMA convolution kernels (rc_conv.hip).
#pragma once

#include <hip/hip_runtime.h>

namespace rc {

constexpr int NET_LAYERS = 6;
constexpr int NET_C = 32;                       // channels of every stored tensor
constexpr int NET_WLAYER = 9 * NET_C * NET_C;   // weights per layer [tap][co][ci]

__device__ __host__ inline unsigned short f2bf(float f) {  // round to nearest even
	unsigned u = __builtin_bit_cast(unsigned, f);
	if((u & 0x7F800000u) == 0x7F800000u) return (unsigned short) ((u >> 16) | ((u & 0xFFFF) ? 0x40 : 0));
	u += 0x7FFFu + ((u >> 16) & 1u);
	return (unsigned short) (u >> 16);
}
__device__ __host__ inline float bf2f(unsigned short h) { return __builtin_bit_cast(float, ((unsigned) h) << 16); }

// run (optional): the launch does nothing when *run == 0 on the device
hipError_t launch_conv3x3(const unsigned short *in, const unsigned short *w, const float *bias, unsigned short *out,
		const unsigned short *mask, float *fout, int N, int H, int W, int mode, hipStream_t s, const int *run = nullptr);
hipError_t launch_conv_wgrad(const unsigned short *in, const unsigned short *g, float *part, int N, int H, int W, hipStream_t s);
int conv_wgrad_blocks(int N, int H);
hipError_t launch_sum_parts(const float *part, float *grad, int nblk, int len, hipStream_t s);
hipError_t launch_adam(float *p, float *m, float *v, const float *g, const float *mask, int len, float lr, float b1,
		float b2, float eps, int t, float scale, hipStream_t s);
hipError_t launch_weights_to_bf16(const float *wf, unsigned short *w, unsigned short *wt, int layers, hipStream_t s);
hipError_t launch_pack_features(const float *x, unsigned short *act, long long npix, hipStream_t s);
hipError_t launch_mse(const float *z, const float *y, unsigned short *g, double *part, long long npix, float inv_count,
		hipStream_t s);

}  // namespace rc
