// rc_net.hip -- the learned CG warm start network behind the C ABI
// (rc_net_* in include/rheocell.h): 6 BF16 conv layers with FP32 master
// weights, trained on the GPU with the kernels of rc_conv.hip (forward,
// backward-data, backward-weights, Adam).  No counterpart in the reference.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/rheocell.h"
#include "rc_conv.h"
#include "rc_net.h"

using namespace rc;

struct rc_net {
	int N = 0, H = 0, W = 0;
	long long npix = 0;
	hipStream_t stream = nullptr;
	float *wf = nullptr, *bf = nullptr;            // fp32 master weights / biases
	float *mw = nullptr, *vw = nullptr, *mb = nullptr, *vb = nullptr;
	float *maskw = nullptr, *maskb = nullptr;
	unsigned short *w = nullptr, *wt = nullptr;    // bf16 [layer][tap][co][ci] and mirrored transpose
	unsigned short *act[NET_LAYERS + 1] = {};      // a0 .. a6 (NHWC32 bf16)
	unsigned short *g[2] = { nullptr, nullptr };   // gradient ping-pong
	float *x = nullptr, *y = nullptr, *out = nullptr;
	float *part = nullptr, *gw = nullptr, *gb = nullptr, *gs = nullptr;
	double *lpart = nullptr;
	int adam_t = 0;
	std::string err;
};

static int nfail(rc_net *n, int code, const char *fmt, ...) {
	char buf[512];
	va_list ap;
	va_start(ap, fmt);
	vsnprintf(buf, sizeof(buf), fmt, ap);
	va_end(ap);
	if(n) n->err = buf;
	return code;
}

#define NCHK(n, call)                                                                                    \
	do {                                                                                                 \
		hipError_t e_ = (call);                                                                          \
		if(e_ != hipSuccess) return nfail(n, RC_ERR_HIP, "%s: %s", #call, hipGetErrorString(e_));        \
	} while(0)

template <typename T>
static hipError_t nalloc(T **p, size_t n) {
	*p = nullptr;
	return hipMalloc((void **) p, (n ? n : 1) * sizeof(T));
}

extern "C" const char *rc_net_last_error(rc_net *n) { return n ? n->err.c_str() : "null net"; }

extern "C" int rc_net_create(int N, int H, int W, rc_net **out) {
	if(!out || N <= 0 || H < 3 || W < 32 || W % 16) return RC_ERR_ARG;
	rc_net *n = new rc_net();
	*out = n;
	n->N = N;
	n->H = H;
	n->W = W;
	n->npix = (long long) N * H * W;
	const size_t nw = NET_LAYERS * NET_WLAYER, nb = NET_LAYERS * NET_C;
	NCHK(n, hipStreamCreateWithFlags(&n->stream, hipStreamNonBlocking));
	NCHK(n, nalloc(&n->wf, nw));
	NCHK(n, nalloc(&n->bf, nb));
	NCHK(n, nalloc(&n->mw, nw));
	NCHK(n, nalloc(&n->vw, nw));
	NCHK(n, nalloc(&n->mb, nb));
	NCHK(n, nalloc(&n->vb, nb));
	NCHK(n, nalloc(&n->maskw, nw));
	NCHK(n, nalloc(&n->maskb, nb));
	NCHK(n, nalloc(&n->w, nw));
	NCHK(n, nalloc(&n->wt, nw));
	NCHK(n, nalloc(&n->gw, nw));
	NCHK(n, nalloc(&n->gb, nb));
	NCHK(n, nalloc(&n->gs, NET_WLAYER + NET_C));
	for(auto &a : n->act) NCHK(n, nalloc(&a, (size_t) n->npix * NET_C));
	for(auto &a : n->g) NCHK(n, nalloc(&a, (size_t) n->npix * NET_C));
	NCHK(n, nalloc(&n->x, (size_t) n->npix * 3));
	NCHK(n, nalloc(&n->y, (size_t) n->npix * 2));
	NCHK(n, nalloc(&n->out, (size_t) n->npix * 2));
	NCHK(n, nalloc(&n->part, (size_t) conv_wgrad_blocks(N, H) * (NET_WLAYER + NET_C)));
	NCHK(n, nalloc(&n->lpart, (size_t) ((n->npix + 255) / 256)));
	// masks: layer 0 reads input channels 0..2, layer 5 writes output channels 0..1
	std::vector<float> mw(nw, 1.f), mb(nb, 1.f);
	for(int l = 0; l < NET_LAYERS; ++l)
		for(int t = 0; t < 9; ++t)
			for(int co = 0; co < NET_C; ++co)
				for(int ci = 0; ci < NET_C; ++ci) {
					const bool ok = (l > 0 || ci < 3) && (l < NET_LAYERS - 1 || co < 2);
					mw[(size_t) l * NET_WLAYER + t * 1024 + co * 32 + ci] = ok ? 1.f : 0.f;
				}
	for(int co = 2; co < NET_C; ++co) mb[(NET_LAYERS - 1) * NET_C + co] = 0.f;
	NCHK(n, hipMemcpy(n->maskw, mw.data(), nw * sizeof(float), hipMemcpyHostToDevice));
	NCHK(n, hipMemcpy(n->maskb, mb.data(), nb * sizeof(float), hipMemcpyHostToDevice));
	NCHK(n, hipMemset(n->wf, 0, nw * sizeof(float)));
	NCHK(n, hipMemset(n->bf, 0, nb * sizeof(float)));
	NCHK(n, hipMemset(n->act[0], 0, (size_t) n->npix * NET_C * sizeof(unsigned short)));
	return RC_OK;
}

extern "C" void rc_net_destroy(rc_net *n) {
	if(!n) return;
	if(n->stream) (void) hipStreamSynchronize(n->stream);
	void *ps[] = { n->wf, n->bf, n->mw, n->vw, n->mb, n->vb, n->maskw, n->maskb, n->w, n->wt, n->x, n->y, n->out,
		n->part, n->gw, n->gb, n->gs, n->lpart, n->g[0], n->g[1] };
	for(void *p : ps)
		if(p) (void) hipFree(p);
	for(auto a : n->act)
		if(a) (void) hipFree(a);
	if(n->stream) (void) hipStreamDestroy(n->stream);
	delete n;
}

extern "C" int rc_net_set_params(rc_net *n, const float *w, const float *b) {
	if(!n || !w || !b) return RC_ERR_ARG;
	const size_t nw = NET_LAYERS * NET_WLAYER, nb = NET_LAYERS * NET_C;
	std::vector<float> ww(w, w + nw), bb(b, b + nb);
	NCHK(n, hipMemcpy(n->wf, ww.data(), nw * sizeof(float), hipMemcpyHostToDevice));
	NCHK(n, hipMemcpy(n->bf, bb.data(), nb * sizeof(float), hipMemcpyHostToDevice));
	NCHK(n, hipMemset(n->mw, 0, nw * sizeof(float)));
	NCHK(n, hipMemset(n->vw, 0, nw * sizeof(float)));
	NCHK(n, hipMemset(n->mb, 0, nb * sizeof(float)));
	NCHK(n, hipMemset(n->vb, 0, nb * sizeof(float)));
	n->adam_t = 0;
	NCHK(n, launch_weights_to_bf16(n->wf, n->w, n->wt, NET_LAYERS, n->stream));
	NCHK(n, hipStreamSynchronize(n->stream));
	return RC_OK;
}

extern "C" int rc_net_get_params(rc_net *n, float *w, float *b) {
	if(!n || !w || !b) return RC_ERR_ARG;
	NCHK(n, hipStreamSynchronize(n->stream));
	NCHK(n, hipMemcpy(w, n->wf, NET_LAYERS * NET_WLAYER * sizeof(float), hipMemcpyDeviceToHost));
	NCHK(n, hipMemcpy(b, n->bf, NET_LAYERS * NET_C * sizeof(float), hipMemcpyDeviceToHost));
	return RC_OK;
}

// forward pass on n->act[0] already packed; result in n->out (fp32, 2 channels);
// the last layer stores no bf16 tensor (backward reads act[0..5] only)
static int forward_packed(rc_net *n, hipStream_t s, const int *run = nullptr) {
	for(int l = 0; l < NET_LAYERS; ++l) {
		const bool last = l == NET_LAYERS - 1;
		NCHK(n, launch_conv3x3(n->act[l], n->w + (size_t) l * NET_WLAYER, n->bf + l * NET_C, last ? nullptr : n->act[l + 1],
				nullptr, last ? n->out : nullptr, n->N, n->H, n->W, last ? 1 : 0, s, run));
	}
	return RC_OK;
}
unsigned short *net_input(rc_net *n) { return n->act[0]; }
int net_forward_packed(rc_net *n, hipStream_t s, const int *run) { return forward_packed(n, s, run); }

int net_forward_device(rc_net *n, const float *d_x, hipStream_t s) {
	NCHK(n, launch_pack_features(d_x, n->act[0], n->npix, s));
	return forward_packed(n, s);
}
const float *net_output(rc_net *n) { return n->out; }
int net_dims(rc_net *n, int *N, int *H, int *W) {
	*N = n->N;
	*H = n->H;
	*W = n->W;
	return 0;
}

extern "C" int rc_net_forward(rc_net *n, const float *x, float *out) {
	if(!n || !x) return RC_ERR_ARG;
	NCHK(n, hipMemcpyAsync(n->x, x, (size_t) n->npix * 3 * sizeof(float), hipMemcpyHostToDevice, n->stream));
	int rc = net_forward_device(n, n->x, n->stream);
	if(rc) return rc;
	if(out) NCHK(n, hipMemcpyAsync(out, n->out, (size_t) n->npix * 2 * sizeof(float), hipMemcpyDeviceToHost, n->stream));
	NCHK(n, hipStreamSynchronize(n->stream));
	return RC_OK;
}

// forward + MSE + backward; gradients in n->gw / n->gb
static int backward(rc_net *n, const float *x, const float *y, double *loss) {
	hipStream_t s = n->stream;
	NCHK(n, hipMemcpyAsync(n->x, x, (size_t) n->npix * 3 * sizeof(float), hipMemcpyHostToDevice, s));
	NCHK(n, hipMemcpyAsync(n->y, y, (size_t) n->npix * 2 * sizeof(float), hipMemcpyHostToDevice, s));
	int rc = net_forward_device(n, n->x, s);
	if(rc) return rc;
	const int lblocks = (int) ((n->npix + 255) / 256);
	NCHK(n, launch_mse(n->out, n->y, n->g[0], n->lpart, n->npix, 1.f / (float) (2 * n->npix), s));
	const int wblk = conv_wgrad_blocks(n->N, n->H);
	int cur = 0;
	for(int l = NET_LAYERS - 1; l >= 0; --l) {
		NCHK(n, launch_conv_wgrad(n->act[l], n->g[cur], n->part, n->N, n->H, n->W, s));
		// sum the partials [blk][tap][co][ci] + [blk][co] into gs, then split into gw[l], gb[l]
		NCHK(n, launch_sum_parts(n->part, n->gs, wblk, NET_WLAYER + NET_C, s));
		NCHK(n, hipMemcpyAsync(n->gw + (size_t) l * NET_WLAYER, n->gs, NET_WLAYER * sizeof(float), hipMemcpyDeviceToDevice, s));
		NCHK(n, hipMemcpyAsync(n->gb + l * NET_C, n->gs + NET_WLAYER, NET_C * sizeof(float), hipMemcpyDeviceToDevice, s));
		if(l > 0)
			NCHK(n, launch_conv3x3(n->g[cur], n->wt + (size_t) l * NET_WLAYER, nullptr, n->g[1 - cur], n->act[l], nullptr,
					n->N, n->H, n->W, 2, s));
		cur = 1 - cur;
	}
	std::vector<double> lp(lblocks);
	NCHK(n, hipMemcpyAsync(lp.data(), n->lpart, lblocks * sizeof(double), hipMemcpyDeviceToHost, s));
	NCHK(n, hipStreamSynchronize(s));
	double L = 0.;
	for(double v : lp) L += v;
	if(loss) *loss = L / (double) (2 * n->npix);
	return RC_OK;
}

extern "C" int rc_net_grads(rc_net *n, const float *x, const float *y, float *gw, float *gb, double *loss) {
	if(!n || !x || !y) return RC_ERR_ARG;
	int rc = backward(n, x, y, loss);
	if(rc) return rc;
	if(gw) NCHK(n, hipMemcpy(gw, n->gw, NET_LAYERS * NET_WLAYER * sizeof(float), hipMemcpyDeviceToHost));
	if(gb) NCHK(n, hipMemcpy(gb, n->gb, NET_LAYERS * NET_C * sizeof(float), hipMemcpyDeviceToHost));
	return RC_OK;
}

extern "C" int rc_net_train_step(rc_net *n, const float *x, const float *y, float lr, double *loss) {
	if(!n || !x || !y) return RC_ERR_ARG;
	int rc = backward(n, x, y, loss);
	if(rc) return rc;
	n->adam_t += 1;
	NCHK(n, launch_adam(n->wf, n->mw, n->vw, n->gw, n->maskw, NET_LAYERS * NET_WLAYER, lr, 0.9f, 0.999f, 1e-8f, n->adam_t,
			1.f, n->stream));
	NCHK(n, launch_adam(n->bf, n->mb, n->vb, n->gb, n->maskb, NET_LAYERS * NET_C, lr, 0.9f, 0.999f, 1e-8f, n->adam_t, 1.f,
			n->stream));
	NCHK(n, launch_weights_to_bf16(n->wf, n->w, n->wt, NET_LAYERS, n->stream));
	NCHK(n, hipStreamSynchronize(n->stream));
	return RC_OK;
}
