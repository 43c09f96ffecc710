/*
 * nn_oracle.h -- CPU restatement (FP64) of the warm-start surrogate network.
 *
 * TEST INFRASTRUCTURE ONLY: the checker of the GPU surrogate kernels.  Only
 * tests, the smoke check and the benchmark's CPU leg may load it.
 *
 * The surrogate has no counterpart in the reference; this file is the
 * definition the GPU implementation is checked against.  Tensors are NCHW,
 * element (b, c, y, x) at ((b*C + c)*H + y)*W + x.  Convolutions are 3x3,
 * stride 1, zero padding 1, cross-correlation:
 *   y[b,o,i,j] = bias[o] + sum_{c,ky,kx} w[((o*Cin + c)*3 + ky)*3 + kx] * x[b,c,i+ky-1,j+kx-1].
 *
 * Network (input 4 x H x W, H and W multiples of 16; every conv but the last
 * is followed by ReLU; pool = 2x2 mean; upadd(lo, skip) = nearest 2x
 * upsampling of lo plus skip):
 *   a1 = conv0(x)      4->16     a2 = conv1(a1)    16->16
 *   a3 = conv2(pool a2) 16->32   a4 = conv3(a3)    32->32
 *   a5 = conv4(pool a4) 32->48   a6 = conv5(a5)    48->48
 *   a7 = conv6(pool a6) 48->64   a8 = conv7(a7)    64->64
 *   a9 = conv8(pool a8) 64->64
 *   a10 = conv9(upadd(a9, a8))   64->48
 *   a11 = conv10(upadd(a10, a6)) 48->32
 *   a12 = conv11(upadd(a11, a4)) 32->16
 *   a13 = conv12(upadd(a12, a2)) 16->16
 *   out = conv13(a13)            16->3   (no ReLU)
 * Parameters: for conv0..conv13 in order, weights [Cout][Cin][3][3] then
 * bias [Cout]; 201 955 values.
 *
 * Loss: masked mean square error over the fluid nodes, the mask being input
 * channel 0:  L = sum_{b,c,p} m[b,p] (out[b,c,p] - t[b,c,p])^2 / (3 sum_{b,p} m[b,p]).
 *
 * Adam (Kingma & Ba 2015, algorithm 1, bias-corrected):
 *   m = b1 m + (1-b1) g;  v = b2 v + (1-b2) g^2;
 *   p -= lr * (m / (1 - b1^t)) / (sqrt(v / (1 - b2^t)) + eps).
 */
#ifndef PORTLBM_NN_ORACLE_H
#define PORTLBM_NN_ORACLE_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORACLE_NN_CONVS = 14, ORACLE_NN_IN = 4, ORACLE_NN_OUT = 3 };

/* Input/output channels of conv k (k < ORACLE_NN_CONVS). */
void oracle_nn_conv_shape(int k, unsigned *cin, unsigned *cout);
size_t oracle_nn_param_count(void);

/* Primitive operators (y, dx, dw, db are written, not accumulated). */
void oracle_conv3x3(const double *x, const double *w, const double *bias, unsigned B, unsigned Cin, unsigned Cout,
                    unsigned H, unsigned W, int relu, double *y);
/* g = gradient w.r.t. the convolution's output before any ReLU. */
void oracle_conv3x3_backward(const double *x, const double *w, const double *g, unsigned B, unsigned Cin,
                             unsigned Cout, unsigned H, unsigned W, double *dx, double *dw, double *db);
void oracle_pool2(const double *x, unsigned B, unsigned C, unsigned H, unsigned W, double *y);
void oracle_upadd(const double *lo, const double *skip, unsigned B, unsigned C, unsigned H, unsigned W, double *z);

/* Whole network.  in: B x 4 x H x W; out: B x 3 x H x W.  Returns 0, or -1
 * when H or W is not a positive multiple of 16. */
int oracle_nn_forward(const double *params, const double *in, unsigned B, unsigned H, unsigned W, double *out);
/* Loss and its gradient w.r.t. every parameter (grad: param_count values).
 * Returns the loss, or NaN for invalid extents or an all-solid batch. */
double oracle_nn_loss_grad(const double *params, const double *in, const double *target, unsigned B, unsigned H,
                           unsigned W, double *grad);

/* One Adam step, t >= 1. */
void oracle_adam(double *p, double *m, double *v, const double *g, size_t n, double lr, double b1, double b2,
                 double eps, unsigned t);

#ifdef __cplusplus
}
#endif

#endif
