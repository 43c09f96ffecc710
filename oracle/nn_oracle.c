/*
 * nn_oracle.c -- FP64 CPU restatement of the warm-start surrogate network
 * (definitions in nn_oracle.h).  TEST INFRASTRUCTURE ONLY.
 *
 * Written for clarity, not speed: direct loops over the defining sums.  The
 * convolution loops are parallel over output elements only, so results do
 * not depend on the thread count.
 */
#include "nn_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

static const unsigned kCin[ORACLE_NN_CONVS] = {4, 16, 16, 32, 32, 48, 48, 64, 64, 64, 48, 32, 16, 16};
static const unsigned kCout[ORACLE_NN_CONVS] = {16, 16, 32, 32, 48, 48, 64, 64, 64, 48, 32, 16, 16, 3};

void oracle_nn_conv_shape(int k, unsigned *cin, unsigned *cout)
{
    *cin = kCin[k];
    *cout = kCout[k];
}

static size_t conv_params(int k) { return (size_t)kCout[k] * kCin[k] * 9u + kCout[k]; }

size_t oracle_nn_param_count(void)
{
    size_t n = 0;
    for (int k = 0; k < ORACLE_NN_CONVS; ++k)
        n += conv_params(k);
    return n;
}

static inline size_t at4(unsigned b, unsigned c, unsigned y, unsigned x, unsigned C, unsigned H, unsigned W)
{
    return (((size_t)b * C + c) * H + y) * W + x;
}

void oracle_conv3x3(const double *x, const double *w, const double *bias, unsigned B, unsigned Cin, unsigned Cout,
                    unsigned H, unsigned W, int relu, double *y)
{
#pragma omp parallel for collapse(2) schedule(static)
    for (unsigned b = 0; b < B; ++b)
        for (unsigned o = 0; o < Cout; ++o)
            for (unsigned i = 0; i < H; ++i)
                for (unsigned j = 0; j < W; ++j)
                {
                    double s = bias ? bias[o] : 0.0;
                    for (unsigned c = 0; c < Cin; ++c)
                        for (int ky = 0; ky < 3; ++ky)
                            for (int kx = 0; kx < 3; ++kx)
                            {
                                const long yy = (long)i + ky - 1, xx = (long)j + kx - 1;
                                if (yy < 0 || xx < 0 || yy >= (long)H || xx >= (long)W)
                                    continue;
                                s += w[(((size_t)o * Cin + c) * 3 + ky) * 3 + kx] *
                                     x[at4(b, c, (unsigned)yy, (unsigned)xx, Cin, H, W)];
                            }
                    y[at4(b, o, i, j, Cout, H, W)] = (relu && !(s > 0.0)) ? 0.0 : s;
                }
}

void oracle_conv3x3_backward(const double *x, const double *w, const double *g, unsigned B, unsigned Cin,
                             unsigned Cout, unsigned H, unsigned W, double *dx, double *dw, double *db)
{
    /* dw[o,c,ky,kx] = sum_{b,i,j} g[b,o,i,j] x[b,c,i+ky-1,j+kx-1];  db[o] = sum g[b,o,.,.] */
#pragma omp parallel for schedule(static)
    for (unsigned o = 0; o < Cout; ++o)
    {
        double sb = 0.0;
        for (unsigned b = 0; b < B; ++b)
            for (unsigned i = 0; i < H; ++i)
                for (unsigned j = 0; j < W; ++j)
                    sb += g[at4(b, o, i, j, Cout, H, W)];
        db[o] = sb;
        for (unsigned c = 0; c < Cin; ++c)
            for (int ky = 0; ky < 3; ++ky)
                for (int kx = 0; kx < 3; ++kx)
                {
                    double s = 0.0;
                    for (unsigned b = 0; b < B; ++b)
                        for (unsigned i = 0; i < H; ++i)
                            for (unsigned j = 0; j < W; ++j)
                            {
                                const long yy = (long)i + ky - 1, xx = (long)j + kx - 1;
                                if (yy < 0 || xx < 0 || yy >= (long)H || xx >= (long)W)
                                    continue;
                                s += g[at4(b, o, i, j, Cout, H, W)] * x[at4(b, c, (unsigned)yy, (unsigned)xx, Cin, H, W)];
                            }
                    dw[(((size_t)o * Cin + c) * 3 + ky) * 3 + kx] = s;
                }
    }
    if (!dx)
        return;
    /* dx[b,c,y,x] = sum_{o,ky,kx} w[o,c,ky,kx] g[b,o,y-ky+1,x-kx+1] */
#pragma omp parallel for collapse(2) schedule(static)
    for (unsigned b = 0; b < B; ++b)
        for (unsigned c = 0; c < Cin; ++c)
            for (unsigned yv = 0; yv < H; ++yv)
                for (unsigned xv = 0; xv < W; ++xv)
                {
                    double s = 0.0;
                    for (unsigned o = 0; o < Cout; ++o)
                        for (int ky = 0; ky < 3; ++ky)
                            for (int kx = 0; kx < 3; ++kx)
                            {
                                const long i = (long)yv - ky + 1, j = (long)xv - kx + 1;
                                if (i < 0 || j < 0 || i >= (long)H || j >= (long)W)
                                    continue;
                                s += w[(((size_t)o * Cin + c) * 3 + ky) * 3 + kx] *
                                     g[at4(b, o, (unsigned)i, (unsigned)j, Cout, H, W)];
                            }
                    dx[at4(b, c, yv, xv, Cin, H, W)] = s;
                }
}

void oracle_pool2(const double *x, unsigned B, unsigned C, unsigned H, unsigned W, double *y)
{
    const unsigned h = H / 2, w = W / 2;
    for (unsigned b = 0; b < B; ++b)
        for (unsigned c = 0; c < C; ++c)
            for (unsigned i = 0; i < h; ++i)
                for (unsigned j = 0; j < w; ++j)
                    y[at4(b, c, i, j, C, h, w)] =
                        0.25 * (x[at4(b, c, 2 * i, 2 * j, C, H, W)] + x[at4(b, c, 2 * i, 2 * j + 1, C, H, W)] +
                                x[at4(b, c, 2 * i + 1, 2 * j, C, H, W)] + x[at4(b, c, 2 * i + 1, 2 * j + 1, C, H, W)]);
}

/* z (B x C x H x W) = nearest 2x upsampling of lo (B x C x H/2 x W/2) + skip */
void oracle_upadd(const double *lo, const double *skip, unsigned B, unsigned C, unsigned H, unsigned W, double *z)
{
    for (unsigned b = 0; b < B; ++b)
        for (unsigned c = 0; c < C; ++c)
            for (unsigned i = 0; i < H; ++i)
                for (unsigned j = 0; j < W; ++j)
                    z[at4(b, c, i, j, C, H, W)] =
                        lo[at4(b, c, i / 2, j / 2, C, H / 2, W / 2)] + skip[at4(b, c, i, j, C, H, W)];
}

/* ---- whole network ---------------------------------------------------- */

typedef struct
{
    const double *w[ORACLE_NN_CONVS], *b[ORACLE_NN_CONVS];
} views;

static void param_views(const double *params, views *v)
{
    size_t off = 0;
    for (int k = 0; k < ORACLE_NN_CONVS; ++k)
    {
        v->w[k] = params + off;
        off += (size_t)kCout[k] * kCin[k] * 9u;
        v->b[k] = params + off;
        off += kCout[k];
    }
}

/* Activations of one forward pass (index names follow nn_oracle.h). */
typedef struct
{
    unsigned B, H[5], W[5];
    double *a[14]; /* a[1..13] */
    double *p[5];  /* p[1..4]: pooled inputs of levels 1..4 */
    double *z[4];  /* z[0..3]: upadd results of levels 0..3 */
    double *out;
} acts;

static double *alloc_t(unsigned B, unsigned C, unsigned H, unsigned W)
{
    return (double *)calloc((size_t)B * C * H * W, sizeof(double));
}

static void acts_free(acts *A)
{
    for (int i = 0; i < 14; ++i)
        free(A->a[i]);
    for (int i = 0; i < 5; ++i)
        free(A->p[i]);
    for (int i = 0; i < 4; ++i)
        free(A->z[i]);
    free(A->out);
}

static int forward_all(const double *params, const double *in, unsigned B, unsigned H, unsigned W, acts *A)
{
    memset(A, 0, sizeof(*A));
    if (B == 0 || H == 0 || W == 0 || H % 16 || W % 16)
        return -1;
    views v;
    param_views(params, &v);
    A->B = B;
    for (int l = 0; l < 5; ++l)
    {
        A->H[l] = H >> l;
        A->W[l] = W >> l;
    }
#define HW(l) A->H[l], A->W[l]
    /* encoder: conv k at level k/2 for k < 8, conv8 at level 4 */
    const double *x = in;
    int k = 0;
    for (int l = 0; l < 4; ++l)
    {
        if (l > 0)
        {
            A->p[l] = alloc_t(B, kCout[k - 1], HW(l));
            oracle_pool2(A->a[k], B, kCout[k - 1], HW(l - 1), A->p[l]);
            x = A->p[l];
        }
        A->a[k + 1] = alloc_t(B, kCout[k], HW(l));
        oracle_conv3x3(x, v.w[k], v.b[k], B, kCin[k], kCout[k], HW(l), 1, A->a[k + 1]);
        ++k;
        A->a[k + 1] = alloc_t(B, kCout[k], HW(l));
        oracle_conv3x3(A->a[k], v.w[k], v.b[k], B, kCin[k], kCout[k], HW(l), 1, A->a[k + 1]);
        ++k;
    }
    /* k == 8: bottleneck */
    A->p[4] = alloc_t(B, kCout[7], HW(4));
    oracle_pool2(A->a[8], B, kCout[7], HW(3), A->p[4]);
    A->a[9] = alloc_t(B, kCout[8], HW(4));
    oracle_conv3x3(A->p[4], v.w[8], v.b[8], B, kCin[8], kCout[8], HW(4), 1, A->a[9]);
    /* decoder: conv 9+m at level 3-m on upadd(a[9+m], skip a[8-2m]) */
    for (int m = 0; m < 4; ++m)
    {
        const int l = 3 - m, kk = 9 + m;
        A->z[l] = alloc_t(B, kCin[kk], HW(l));
        oracle_upadd(A->a[9 + m], A->a[8 - 2 * m], B, kCin[kk], HW(l), A->z[l]);
        A->a[10 + m] = alloc_t(B, kCout[kk], HW(l));
        oracle_conv3x3(A->z[l], v.w[kk], v.b[kk], B, kCin[kk], kCout[kk], HW(l), 1, A->a[10 + m]);
    }
    A->out = alloc_t(B, kCout[13], HW(0));
    oracle_conv3x3(A->a[13], v.w[13], v.b[13], B, kCin[13], kCout[13], HW(0), 0, A->out);
#undef HW
    return 0;
}

int oracle_nn_forward(const double *params, const double *in, unsigned B, unsigned H, unsigned W, double *out)
{
    acts A;
    if (forward_all(params, in, B, H, W, &A))
    {
        acts_free(&A);
        return -1;
    }
    memcpy(out, A.out, (size_t)B * ORACLE_NN_OUT * H * W * sizeof(double));
    acts_free(&A);
    return 0;
}

/* g := g * [a > 0] */
static void relu_mask(double *g, const double *a, size_t n)
{
    for (size_t i = 0; i < n; ++i)
        if (!(a[i] > 0.0))
            g[i] = 0.0;
}

/* dx += pool backward of dy (dx: B x C x H x W, dy: half size) */
static void pool2_backward_add(const double *dy, unsigned B, unsigned C, unsigned H, unsigned W, double *dx)
{
    for (unsigned b = 0; b < B; ++b)
        for (unsigned c = 0; c < C; ++c)
            for (unsigned i = 0; i < H; ++i)
                for (unsigned j = 0; j < W; ++j)
                    dx[at4(b, c, i, j, C, H, W)] += 0.25 * dy[at4(b, c, i / 2, j / 2, C, H / 2, W / 2)];
}

/* dlo = sum of dz over each 2x2 block (dlo: half size) */
static void upsample_backward(const double *dz, unsigned B, unsigned C, unsigned H, unsigned W, double *dlo)
{
    const unsigned h = H / 2, w = W / 2;
    for (unsigned b = 0; b < B; ++b)
        for (unsigned c = 0; c < C; ++c)
            for (unsigned i = 0; i < h; ++i)
                for (unsigned j = 0; j < w; ++j)
                    dlo[at4(b, c, i, j, C, h, w)] =
                        dz[at4(b, c, 2 * i, 2 * j, C, H, W)] + dz[at4(b, c, 2 * i, 2 * j + 1, C, H, W)] +
                        dz[at4(b, c, 2 * i + 1, 2 * j, C, H, W)] + dz[at4(b, c, 2 * i + 1, 2 * j + 1, C, H, W)];
}

double oracle_nn_loss_grad(const double *params, const double *in, const double *target, unsigned B, unsigned H,
                           unsigned W, double *grad)
{
    acts A;
    if (forward_all(params, in, B, H, W, &A))
    {
        acts_free(&A);
        return NAN;
    }
    const size_t px = (size_t)H * W;
    double msum = 0.0;
    for (unsigned b = 0; b < B; ++b)
        for (size_t p = 0; p < px; ++p)
            msum += in[((size_t)b * ORACLE_NN_IN + 0) * px + p];
    if (!(msum > 0.0))
    {
        acts_free(&A);
        return NAN;
    }
    const double norm = 1.0 / (ORACLE_NN_OUT * msum);
    double loss = 0.0;
    double *dout = alloc_t(B, ORACLE_NN_OUT, H, W);
    for (unsigned b = 0; b < B; ++b)
        for (unsigned c = 0; c < ORACLE_NN_OUT; ++c)
            for (size_t p = 0; p < px; ++p)
            {
                const double m = in[((size_t)b * ORACLE_NN_IN + 0) * px + p];
                const size_t i = ((size_t)b * ORACLE_NN_OUT + c) * px + p;
                const double e = A.out[i] - target[i];
                loss += m * e * e;
                dout[i] = 2.0 * m * e * norm;
            }
    loss *= norm;

    /* gradient buffers: dw/db views into grad */
    double *gw[ORACLE_NN_CONVS], *gb[ORACLE_NN_CONVS];
    {
        size_t off = 0;
        for (int k = 0; k < ORACLE_NN_CONVS; ++k)
        {
            gw[k] = grad + off;
            off += (size_t)kCout[k] * kCin[k] * 9u;
            gb[k] = grad + off;
            off += kCout[k];
        }
    }
    views v;
    param_views(params, &v);
    double *da[14] = {0};
#define HW(l) A.H[l], A.W[l]
#define NEL(c, l) ((size_t)B * (c) * A.H[l] * A.W[l])
    /* conv13 (no ReLU): input a13 */
    da[13] = alloc_t(B, kCin[13], HW(0));
    oracle_conv3x3_backward(A.a[13], v.w[13], dout, B, kCin[13], kCout[13], HW(0), da[13], gw[13], gb[13]);
    /* decoder, m = 3..0: conv 9+m at level 3-m, output a[10+m], input z[3-m] */
    for (int m = 3; m >= 0; --m)
    {
        const int l = 3 - m, kk = 9 + m;
        relu_mask(da[10 + m], A.a[10 + m], NEL(kCout[kk], l));
        double *dz = alloc_t(B, kCin[kk], HW(l));
        oracle_conv3x3_backward(A.z[l], v.w[kk], da[10 + m], B, kCin[kk], kCout[kk], HW(l), dz, gw[kk], gb[kk]);
        /* z = up(a[9+m]) + a[8-2m] */
        da[9 + m] = alloc_t(B, kCin[kk], HW(l + 1));
        upsample_backward(dz, B, kCin[kk], HW(l), da[9 + m]);
        da[8 - 2 * m] = dz; /* the skip's gradient so far */
    }
    /* bottleneck conv8: output a9, input p4 = pool(a8) */
    relu_mask(da[9], A.a[9], NEL(kCout[8], 4));
    {
        double *dp = alloc_t(B, kCin[8], HW(4));
        oracle_conv3x3_backward(A.p[4], v.w[8], da[9], B, kCin[8], kCout[8], HW(4), dp, gw[8], gb[8]);
        pool2_backward_add(dp, B, kCin[8], HW(3), da[8]);
        free(dp);
    }
    /* encoder, level l = 3..0: conv 2l+1 (a[2l+1] -> a[2l+2]), conv 2l (input p[l] or in -> a[2l+1]) */
    for (int l = 3; l >= 0; --l)
    {
        const int k1 = 2 * l + 1, k0 = 2 * l;
        relu_mask(da[k1 + 1], A.a[k1 + 1], NEL(kCout[k1], l));
        da[k1] = alloc_t(B, kCin[k1], HW(l));
        oracle_conv3x3_backward(A.a[k1], v.w[k1], da[k1 + 1], B, kCin[k1], kCout[k1], HW(l), da[k1], gw[k1], gb[k1]);
        relu_mask(da[k1], A.a[k1], NEL(kCout[k0], l));
        if (l == 0)
        {
            oracle_conv3x3_backward(in, v.w[k0], da[k1], B, kCin[k0], kCout[k0], HW(0), NULL, gw[k0], gb[k0]);
        }
        else
        {
            double *dp = alloc_t(B, kCin[k0], HW(l));
            oracle_conv3x3_backward(A.p[l], v.w[k0], da[k1], B, kCin[k0], kCout[k0], HW(l), dp, gw[k0], gb[k0]);
            pool2_backward_add(dp, B, kCin[k0], HW(l - 1), da[k0]);
            free(dp);
        }
    }
#undef HW
#undef NEL
    for (int i = 0; i < 14; ++i)
        free(da[i]);
    free(dout);
    acts_free(&A);
    return loss;
}

void oracle_adam(double *p, double *m, double *v, const double *g, size_t n, double lr, double b1, double b2,
                 double eps, unsigned t)
{
    const double c1 = 1.0 - pow(b1, (double)t), c2 = 1.0 - pow(b2, (double)t);
    for (size_t i = 0; i < n; ++i)
    {
        m[i] = b1 * m[i] + (1.0 - b1) * g[i];
        v[i] = b2 * v[i] + (1.0 - b2) * g[i] * g[i];
        p[i] -= lr * (m[i] / c1) / (sqrt(v[i] / c2) + eps);
    }
}
