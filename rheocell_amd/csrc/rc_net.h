// rc_net.h -- device-side entry points of the learned warm-start network used
// by the CG (rc_api.hip).
#pragma once

#include <hip/hip_runtime.h>

struct rc_net;

// packs d_x (fp32 [N*H*W][3] features) and runs the forward pass on stream s
int net_forward_device(rc_net *n, const float *d_x, hipStream_t s);
// fp32 [N*H*W][2] output of the last forward pass
const float *net_output(rc_net *n);
int net_dims(rc_net *n, int *N, int *H, int *W);
