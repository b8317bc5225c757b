// rc_net.h -- device-side entry points of the learned warm-start network used
// by the CG (rc_api.hip).
#pragma once

#include <hip/hip_runtime.h>

struct rc_net;

// packs d_x (fp32 [N*H*W][3] features) and runs the forward pass on stream s
int net_forward_device(rc_net *n, const float *d_x, hipStream_t s);
// input tensor of the first layer, bf16 NHWC with 32 channels; channels 8..31
// are zero and stay zero, so a producer may write channels 0..7 (16 bytes)
unsigned short *net_input(rc_net *n);
// forward pass on the input tensor as it is; with run set, the layers do
// nothing while *run == 0 on the device
int net_forward_packed(rc_net *n, hipStream_t s, const int *run = nullptr);
// fp32 [N*H*W][2] output of the last forward pass
const float *net_output(rc_net *n);
int net_dims(rc_net *n, int *N, int *H, int *W);
