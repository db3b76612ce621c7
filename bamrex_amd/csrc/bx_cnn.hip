// Learned fine-level smoother for the IMEX pressure Helmholtz solve: a small
// convolutional network, hand-written forward (BF16 MFMA, FP32 accumulate)
// and backward kernels, and its training loop.
//
// Network N: 4 bias-free 3x3 convolutions, 1 -> 16 -> 16 -> 16 -> 1
// channels, ReLU after the first three.  Bias-free ReLU layers make N
// positively homogeneous (N(c x) = c N(x), c > 0) up to BF16 rounding, so the
// correction scales with the residual.  Weights and activations are BF16,
// every sum is accumulated in FP32.  The correction applied to a residual r is
//     e = omega * (sigma / D) * N(r / sigma),   sigma = rms(r),
// with D the diagonal of the Helmholtz operator: N = identity is a Jacobi
// step, the network learns a better local approximate inverse.
//
// Layouts: activations NHWC with 16 channels (bf16, 32 B per pixel).
// Parameters (FP32 master copy, flat, BX_CNN_NPARAMS):
//   W1[9][16]        tap t, out n
//   W2[9][16][16]    tap t, out n, in k
//   W3[9][16][16]
//   W4[9][16]        tap t, in k
// tap t = (dy+1)*3 + (dx+1), neighbour (x+dx, y+dy).
//
// All layers run on v_mfma_f32_16x16x16bf16_1k with the output channels as
// rows and 16 pixels as columns: the 16 -> 16 layers accumulate 9 MFMAs
// (one per tap, K = 16 input channels), the first layer takes the 9 taps as
// K, the last one forms 9 per-tap partials per pixel.  The weight gradients
// are FP32 GEMMs over pixels on v_mfma_f32_16x16x4f32.
//
// Training loss on crops of (residual, error) pairs harvested from plain
// multigrid solves (A e = r):
//     L = sum_crops  ||mask (r - A e_cnn)||^2 / max(||mask r||^2, |mask| sigma^2)
// (mask = crop interior at distance >= 5 from the crop edge, where the 9x9
// receptive field sees no padding, sigma = rms over the crop), i.e. the
// A^2-norm of the remaining error e - e_cnn.  The floor keeps a crop whose
// residual sits in its border (seen by the network, outside the mask) from
// dividing by a near-zero norm: such crops made single-batch spikes of the
// loss of 1e9 and more.  Gradients: dL/de_cnn = -2 A^T (mask res) (A symmetric for PEEK),
// then back through the convolutions; Adam on the FP32 master weights.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "../../include/bamrex_hip.h"
#include "bx_helmholtz.h"

namespace {

typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));
typedef float v2f __attribute__((ext_vector_type(2)));
typedef __bf16 v2bf __attribute__((ext_vector_type(2)));

constexpr int C = 16;
constexpr int NW1 = 9 * C, NW2 = 9 * C * C, NW3 = 9 * C * C, NW4 = 9 * C;
constexpr int OFF1 = 0, OFF2 = NW1, OFF3 = OFF2 + NW2, OFF4 = OFF3 + NW3;
constexpr int NPAR = OFF4 + NW4;
static_assert(NPAR == BX_CNN_NPARAMS, "parameter count");
// training reductions: pixels per weight-gradient block, the row of one
// block's partials (W3 | W2 | W4 | W1), blocks of the loss sum
constexpr int WG_CH       = 256;
constexpr int WG_ROW      = 2 * NW2 + NW4 + NW1;
constexpr int LOSS_BLOCKS = 2048;

__host__ __device__ inline unsigned short f2bf(float f)
{
    unsigned int u = __builtin_bit_cast(unsigned int, f);
    u += 0x7fffu + ((u >> 16) & 1u); // round to nearest even
    return (unsigned short)(u >> 16);
}
__host__ __device__ inline float bf2f(unsigned short h)
{
    return __builtin_bit_cast(float, (unsigned int)h << 16);
}

// Field geometry of one launch: `nimg` images of w x h pixels stored one
// after the other.  PERIODIC: neighbours wrap; otherwise zero padding.
struct Geo
{
    int w, h, nimg;
};

template <bool PERIODIC>
__device__ inline int nbr(int x, int y, int dx, int dy, const Geo &g, int img)
{
    int xx = x + dx, yy = y + dy;
    if (PERIODIC)
    {
        xx = xx < 0 ? xx + g.w : (xx >= g.w ? xx - g.w : xx);
        yy = yy < 0 ? yy + g.h : (yy >= g.h ? yy - g.h : yy);
    }
    else if (xx < 0 || yy < 0 || xx >= g.w || yy >= g.h)
        return -1;
    return (img * g.h + yy) * g.w + xx;
}

// ---- inference / training forward --------------------------------------
// Every layer puts output channels (or, in layer 4, taps) in the MFMA's
// rows and 16 pixels in its columns, so a lane's result is 4 consecutive
// channels of one pixel (one 8-byte store).  v_mfma_f32_16x16x16bf16_1k:
// A[i][k] (lane l: i = l%16, k = 4(l/16)+v), B[k][j] (lane l: j = l%16,
// k = 4(l/16)+v), D[i][j] (lane l: i = 4(l/16)+v, j = l%16).
// Each D[i][j] is a dot product over k of row i of A with column j of B, so a
// pixel's result does not depend on which pixels share its tile, and each
// 16 -> 16 layer adds its 9 taps' products in tap order: the fused inference
// kernel below and these layer-by-layer kernels agree bit for bit.
//   layer 1: A = W1 (out channel x tap), B = the 9 taps of x (k = tap, 7 zero)
//   layers 2, 3: A = W[tap] (out x in channel), B = the neighbour's channels
//   layer 4: A = W4 (tap x in channel), B = h3, giving the per-tap partials
//            P[pixel][tap]; y = sum over taps of P at the neighbours, in tap
//            order.

// ReLU on the bits (a negative float, -0 included, becomes +0; the integer
// maximum is one instruction) and round-to-nearest-even to bf16, two values
// per hardware conversion
__device__ inline unsigned relu_pack2(float lo, float hi)
{
    const int il = __builtin_bit_cast(int, lo), ih = __builtin_bit_cast(int, hi);
    v2f v = { __builtin_bit_cast(float, il > 0 ? il : 0), __builtin_bit_cast(float, ih > 0 ? ih : 0) };
    return __builtin_bit_cast(unsigned, __builtin_convertvector(v, v2bf));
}
// four ReLU'd channels as bf16, packed for one 8-byte store
__device__ inline uint2 relu_pack4(const v4f &acc)
{
    return make_uint2(relu_pack2(acc[0], acc[1]), relu_pack2(acc[2], acc[3]));
}

// A operand of layer 1: A[i = out][k = tap] = W1[tap][out] (tap < 9)
__device__ inline v4s w1_operand(const unsigned short *__restrict__ Wb1, int m, int kg)
{
    v4s b;
#pragma unroll
    for (int v = 0; v < 4; ++v)
    {
        const int t = 4 * kg + v;
        b[v] = t < 9 ? (short)Wb1[t * C + m] : (short)0;
    }
    return b;
}
// A operand of layer 4: A[i = tap][k = in channel] = W4[tap][in] (tap < 9)
__device__ inline v4s w4_operand(const unsigned short *__restrict__ Wb4, int m, int kg)
{
    v4s b;
#pragma unroll
    for (int v = 0; v < 4; ++v)
        b[v] = m < 9 ? (short)Wb4[m * C + 4 * kg + v] : (short)0;
    return b;
}

// layer 1: x (bf16, 1 channel) -> h1 (bf16, 16, ReLU); B = the taps of
// pixel p0 + m (lane: taps 4kg .. 4kg+3)
template <bool PERIODIC>
__global__ __launch_bounds__(256) void k_conv1(const unsigned short *__restrict__ xin,
                                               unsigned short *__restrict__ hout,
                                               const unsigned short *__restrict__ Wb1,
                                               Geo g)
{
    const int       lane = threadIdx.x & 63;
    const long long tile = ((long long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const long long npx  = (long long)g.w * g.h * g.nimg;
    const long long p0   = tile * 16;
    if (p0 >= npx) // whole wave exits together (npx is a multiple of 16)
        return;
    const unsigned wh_ = (unsigned)g.w * (unsigned)g.h;
    const int img = (int)((unsigned)p0 / wh_);
    const int rem = (int)((unsigned)p0 - (unsigned)img * wh_);
    const int y = (int)((unsigned)rem / (unsigned)g.w), x0 = rem - y * g.w;
    const int m = lane & 15, kg = lane >> 4;
    v4s       a;
#pragma unroll
    for (int v = 0; v < 4; ++v)
    {
        const int t = 4 * kg + v;
        short     xv = 0;
        if (t < 9)
        {
            const int q = nbr<PERIODIC>(x0 + m, y, t % 3 - 1, t / 3 - 1, g, img);
            if (q >= 0)
                xv = (short)xin[q];
        }
        a[v] = xv;
    }
    v4f acc = { 0.f, 0.f, 0.f, 0.f };
    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(w1_operand(Wb1, m, kg), a, acc, 0, 0, 0);
    // lane: channels 4kg .. 4kg+3 of pixel p0 + m
    *reinterpret_cast<uint2 *>(hout + (p0 + m) * C + 4 * kg) = relu_pack4(acc);
}

// 16 -> 16 layers, D[i = out channel][j = pixel]: 9 MFMAs in tap order,
// A = W[tap] (lane l: out channel l%16, input channels 4(l/16) .. +3),
// B = the neighbour's channels (lane l: pixel l%16, the same 4 channels).
// Wb: bf16 weights [9][16 out][16 in].
__device__ inline v4s w16_tap(const unsigned short *__restrict__ Wb, int t, int i, int kg)
{
    return *reinterpret_cast<const v4s *>(Wb + (t * C + i) * C + 4 * kg);
}

// layers 2, 3: h (bf16, 16) -> h' (bf16, 16, ReLU), one wave per 16 pixels
template <bool PERIODIC>
__global__ __launch_bounds__(256) void k_conv16(const unsigned short *__restrict__ hin,
                                                unsigned short *__restrict__ hout,
                                                const unsigned short *__restrict__ Wb,
                                                Geo g)
{
    const int       lane  = threadIdx.x & 63;
    const long long tile  = ((long long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const long long npx   = (long long)g.w * g.h * g.nimg;
    const long long p0    = tile * 16;
    if (p0 >= npx) // whole wave exits together (npx is a multiple of 16)
        return;
    const unsigned wh_ = (unsigned)g.w * (unsigned)g.h;
    const int img = (int)((unsigned)p0 / wh_);
    const int rem = (int)((unsigned)p0 - (unsigned)img * wh_);
    const int y = (int)((unsigned)rem / (unsigned)g.w), x0 = rem - y * g.w;
    const int m = lane & 15, kg = lane >> 4;
    v4f       acc = { 0.f, 0.f, 0.f, 0.f };
#pragma unroll
    for (int t = 0; t < 9; ++t)
    {
        const int q = nbr<PERIODIC>(x0 + m, y, t % 3 - 1, t / 3 - 1, g, img);
        v4s       b = { 0, 0, 0, 0 };
        if (q >= 0)
            b = *reinterpret_cast<const v4s *>(hin + (long long)q * C + 4 * kg);
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(w16_tap(Wb, t, m, kg), b, acc, 0, 0, 0);
    }
    // lane: channels 4kg .. 4kg+3 of pixel p0 + m
    *reinterpret_cast<uint2 *>(hout + (p0 + m) * C + 4 * kg) = relu_pack4(acc);
}

// layer 4, part 1: per-tap partials P[p][t] = sum_k W4[t][k] h3[p][k]
__global__ __launch_bounds__(256) void k_conv4_taps(const unsigned short *__restrict__ hin,
                                                    float *__restrict__ P,
                                                    const unsigned short *__restrict__ Wb4,
                                                    long long npx)
{
    const int       lane = threadIdx.x & 63;
    const long long tile = ((long long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const long long p0   = tile * 16;
    if (p0 >= npx)
        return;
    const int m = lane & 15, kg = lane >> 4;
    const v4s b = *reinterpret_cast<const v4s *>(hin + (p0 + m) * C + 4 * kg);
    v4f       acc = { 0.f, 0.f, 0.f, 0.f };
    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(w4_operand(Wb4, m, kg), b, acc, 0, 0, 0);
    // lane: taps 4kg .. 4kg+3 of pixel p0 + m
    *reinterpret_cast<v4f *>(P + (p0 + m) * C + 4 * kg) = acc;
}

// layer 4, part 2: y[p] = sum over taps t (in order) of P[p + off_t][t]
template <bool PERIODIC>
__global__ __launch_bounds__(256) void k_conv4_sum(const float *__restrict__ P,
                                                   float *__restrict__ yout, Geo g)
{
    const long long npx = (long long)g.w * g.h * g.nimg;
    const long long p   = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= npx)
        return;
    const unsigned wh_ = (unsigned)g.w * (unsigned)g.h;
    const int img = (int)((unsigned)p / wh_);
    const int rem = (int)((unsigned)p - (unsigned)img * wh_);
    const int y = (int)((unsigned)rem / (unsigned)g.w), x = rem - y * g.w;
    float     s = 0.f;
#pragma unroll
    for (int t = 0; t < 9; ++t)
    {
        const int q = nbr<PERIODIC>(x, y, t % 3 - 1, t / 3 - 1, g, img);
        if (q < 0)
            continue;
        s += P[(long long)q * C + t];
    }
    yout[p] = s;
}



// sum of squares of r, one partial per block (the fused kernel adds the
// partials in a fixed order)
constexpr int SS_BLOCKS = 2048;
__global__ __launch_bounds__(256) void k_sumsq(const double *__restrict__ r,
                                               long long n,
                                               double *__restrict__ partial)
{
    __shared__ double sh[256];
    double            s = 0;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        s += r[i] * r[i];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1)
    {
        if ((int)threadIdx.x < off)
            sh[threadIdx.x] += sh[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        partial[blockIdx.x] = sh[0];
}

// operator diagonal: alpha a + dhx (bXl + bXr) + dhy (bYl + bYr)
__global__ __launch_bounds__(256) void k_diag(const double *__restrict__ a,
                                              const double *__restrict__ bX,
                                              const double *__restrict__ bY,
                                              double *__restrict__ D, int nx, int ny,
                                              double alpha, double dhx, double dhy)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
    if (i >= nx)
        return;
    const long long c = (long long)j * nx + i;
    D[c] = alpha * a[c] + dhx * (bX[(long long)j * (nx + 1) + i] + bX[(long long)j * (nx + 1) + i + 1])
           + dhy * (bY[(long long)j * nx + i] + bY[(long long)(j + 1) * nx + i]);
}

// ---- training: crop gather, loss, backward -------------------------------
struct CropSet
{
    const double *R;  // fields [nf][ny][nx]
    const int    *xy; // per crop: field, x0, y0
    int           nx, ny, cs; // field size, crop size
};

__device__ inline long long gidx(const CropSet &c, int f, int x, int y)
{
    x = x % c.nx;
    y = y % c.ny;
    if (x < 0) x += c.nx;
    if (y < 0) y += c.ny;
    return ((long long)f * c.ny + y) * c.nx + x;
}

// per crop: sigma (rms over the crop) and the loss normaliser
// max(||mask r||^2, |mask| sigma^2): floored, since a crop whose residual
// sits in its border (which the network sees) has a near-zero masked norm
__global__ void k_crop_stats(CropSet c, double *__restrict__ sig,
                             double *__restrict__ rn2, int border)
{
    const int b = blockIdx.x;
    const int f = c.xy[3 * b], x0 = c.xy[3 * b + 1], y0 = c.xy[3 * b + 2];
    __shared__ double s1[256], s2[256];
    double            a1 = 0, a2 = 0;
    for (int p = threadIdx.x; p < c.cs * c.cs; p += blockDim.x)
    {
        const int    y = p / c.cs, x = p - y * c.cs;
        const double v = c.R[gidx(c, f, x0 + x, y0 + y)];
        a1 += v * v;
        if (x >= border && y >= border && x < c.cs - border && y < c.cs - border)
            a2 += v * v;
    }
    s1[threadIdx.x] = a1;
    s2[threadIdx.x] = a2;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1)
    {
        if ((int)threadIdx.x < off)
        {
            s1[threadIdx.x] += s1[threadIdx.x + off];
            s2[threadIdx.x] += s2[threadIdx.x + off];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0)
    {
        const int    in    = c.cs - 2 * border;
        const double floor = s1[0] * ((double)(in * in) / (double)(c.cs * c.cs));
        sig[b]             = sqrt(s1[0] / (double)(c.cs * c.cs));
        rn2[b]             = s2[0] < floor ? floor : s2[0];
    }
}

__global__ __launch_bounds__(256) void k_crop_prep(CropSet c, const double *__restrict__ sig,
                                                   unsigned short *__restrict__ x,
                                                   int batch)
{
    const long long n = (long long)batch * c.cs * c.cs;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
    {
        const int b = (int)(i / (c.cs * c.cs));
        const int p = (int)(i - (long long)b * c.cs * c.cs);
        const int y = p / c.cs, xx = p - y * c.cs;
        const int f = c.xy[3 * b], x0 = c.xy[3 * b + 1], y0 = c.xy[3 * b + 2];
        const double s = sig[b] > 0 ? 1.0 / sig[b] : 0.0;
        x[i] = f2bf((float)(c.R[gidx(c, f, x0 + xx, y0 + y)] * s));
    }
}

struct OpRef
{
    const double *a, *bX, *bY, *D;
    int           nx, ny;
    double        alpha, dhx, dhy;
};

// A applied at global cell (gx, gy) to a crop-local field u (zero outside)
__device__ inline double apply_local(const OpRef &o, const double *u, int cs,
                                     int x, int y, int gx, int gy)
{
    auto at = [&](int xx, int yy) -> double {
        return (xx < 0 || yy < 0 || xx >= cs || yy >= cs) ? 0.0 : u[yy * cs + xx];
    };
    gx %= o.nx; gy %= o.ny;
    if (gx < 0) gx += o.nx;
    if (gy < 0) gy += o.ny;
    const long long c   = (long long)gy * o.nx + gx;
    const double    bxl = o.bX[(long long)gy * (o.nx + 1) + gx];
    const double    bxr = o.bX[(long long)gy * (o.nx + 1) + gx + 1];
    const double    byl = o.bY[(long long)gy * o.nx + gx];
    const double    byr = o.bY[(long long)(gy + 1) * o.nx + gx];
    const double    uc  = at(x, y);
    return o.alpha * o.a[c] * uc
           - o.dhx * (bxr * (at(x + 1, y) - uc) - bxl * (uc - at(x - 1, y)))
           - o.dhy * (byr * (at(x, y + 1) - uc) - byl * (uc - at(x, y - 1)));
}

__device__ inline double Dat(const OpRef &o, int gx, int gy)
{
    gx %= o.nx; gy %= o.ny;
    if (gx < 0) gx += o.nx;
    if (gy < 0) gy += o.ny;
    return o.D[(long long)gy * o.nx + gx];
}

// e_cnn = sigma * y / D for every crop pixel
__global__ __launch_bounds__(256) void k_crop_e(CropSet c, OpRef o,
                                                const double *__restrict__ sig,
                                                const float *__restrict__ y,
                                                double *__restrict__ e, int batch)
{
    const long long n = (long long)batch * c.cs * c.cs;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
    {
        const int b = (int)(i / (c.cs * c.cs));
        const int p = (int)(i - (long long)b * c.cs * c.cs);
        const int yy = p / c.cs, xx = p - yy * c.cs;
        e[i] = sig[b] * (double)y[i] / Dat(o, c.xy[3 * b + 1] + xx, c.xy[3 * b + 2] + yy);
    }
}

// res = mask (r - A e); loss contribution res^2 / ||mask r||^2, one partial
// sum per block (summed in a fixed order by k_sum_partials_f64)
__global__ __launch_bounds__(256) void k_crop_res(CropSet c, OpRef o,
                                                  const double *__restrict__ e,
                                                  const double *__restrict__ rn2,
                                                  double *__restrict__ res,
                                                  double *__restrict__ loss_part,
                                                  int batch, int border)
{
    const long long n = (long long)batch * c.cs * c.cs;
    double          acc = 0;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
    {
        const int b = (int)(i / (c.cs * c.cs));
        const int p = (int)(i - (long long)b * c.cs * c.cs);
        const int yy = p / c.cs, xx = p - yy * c.cs;
        double    v  = 0;
        if (xx >= border && yy >= border && xx < c.cs - border && yy < c.cs - border)
        {
            const int f = c.xy[3 * b], gx = c.xy[3 * b + 1] + xx,
                      gy = c.xy[3 * b + 2] + yy;
            v = c.R[gidx(c, f, gx, gy)]
                - apply_local(o, e + (long long)b * c.cs * c.cs, c.cs, xx, yy, gx, gy);
            acc += v * v / (rn2[b] > 0 ? rn2[b] : 1.0);
        }
        res[i] = v;
    }
    __shared__ double red[4];
    for (int off = 32; off > 0; off >>= 1)
        acc += __shfl_xor(acc, off);
    if ((threadIdx.x & 63) == 0)
        red[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0)
        loss_part[blockIdx.x] = (red[0] + red[1]) + (red[2] + red[3]);
}

// *out = sum of n partials in a fixed order: thread t of one 256-thread
// block adds a contiguous run of them in index order, then a fixed tree
// adds the 256 run sums
__global__ __launch_bounds__(256) void k_sum_partials_f64(const double *__restrict__ part, int n,
                                                          double *__restrict__ out)
{
    __shared__ double sh[256];
    const int per = (n + 255) / 256, b0 = threadIdx.x * per, b1 = min(n, b0 + per);
    double    s = 0;
    for (int b = b0; b < b1; ++b)
        s += part[b];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1)
    {
        if ((int)threadIdx.x < off)
            sh[threadIdx.x] += sh[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        *out = sh[0];
}

// g_y = dL/dy = -2 A^T(res / ||mask r||^2) * sigma / D   (A symmetric)
__global__ __launch_bounds__(256) void k_crop_gy(CropSet c, OpRef o,
                                                 const double *__restrict__ res,
                                                 const double *__restrict__ rn2,
                                                 const double *__restrict__ sig,
                                                 float *__restrict__ gy, int batch,
                                                 float scale)
{
    const long long n = (long long)batch * c.cs * c.cs;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
    {
        const int b = (int)(i / (c.cs * c.cs));
        const int p = (int)(i - (long long)b * c.cs * c.cs);
        const int yy = p / c.cs, xx = p - yy * c.cs;
        const int gx = c.xy[3 * b + 1] + xx, gyy = c.xy[3 * b + 2] + yy;
        const double ar = apply_local(o, res + (long long)b * c.cs * c.cs, c.cs, xx,
                                      yy, gx, gyy);
        const double ge = -2.0 * ar / (rn2[b] > 0 ? rn2[b] : 1.0);
        gy[i] = (float)(scale * ge * sig[b] / Dat(o, gx, gyy));
    }
}

// layer 4 backward: g_h3 = relu'(h3) * sum_t W4[t] g_y[p - off_t]
__global__ __launch_bounds__(256) void k_bwd4_data(const float *__restrict__ gy,
                                                   const unsigned short *__restrict__ h3,
                                                   const float *__restrict__ W4,
                                                   float *__restrict__ gh, Geo g)
{
    const long long npx = (long long)g.w * g.h * g.nimg;
    const long long p   = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= npx)
        return;
    const unsigned wh_ = (unsigned)g.w * (unsigned)g.h;
    const int img = (int)((unsigned)p / wh_);
    const int rem = (int)((unsigned)p - (unsigned)img * wh_);
    const int y = (int)((unsigned)rem / (unsigned)g.w), x = rem - y * g.w;
    float     acc[C];
#pragma unroll
    for (int k = 0; k < C; ++k)
        acc[k] = 0.f;
#pragma unroll
    for (int t = 0; t < 9; ++t)
    {
        // output pixel o with o + off_t = p  ->  o = p - off_t
        const int q = nbr<false>(x, y, -(t % 3 - 1), -(t / 3 - 1), g, img);
        if (q < 0)
            continue;
        const float gv = gy[q];
#pragma unroll
        for (int k = 0; k < C; ++k)
            acc[k] += W4[t * C + k] * gv;
    }
#pragma unroll
    for (int k = 0; k < C; ++k)
        gh[p * C + k] = bf2f(h3[p * C + k]) > 0.f ? acc[k] : 0.f;
}

// 16->16 layer backward (data): g_in[p][k] = relu'(h_in) sum_t sum_n W[t][n][k] g_out[p - off_t][n]
__global__ __launch_bounds__(256) void k_bwd16_data(const float *__restrict__ gout,
                                                    const unsigned short *__restrict__ hin,
                                                    const float *__restrict__ W,
                                                    float *__restrict__ gin, Geo g)
{
    const long long npx = (long long)g.w * g.h * g.nimg;
    const long long p   = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= npx)
        return;
    const unsigned wh_ = (unsigned)g.w * (unsigned)g.h;
    const int img = (int)((unsigned)p / wh_);
    const int rem = (int)((unsigned)p - (unsigned)img * wh_);
    const int y = (int)((unsigned)rem / (unsigned)g.w), x = rem - y * g.w;
    float     acc[C];
#pragma unroll
    for (int k = 0; k < C; ++k)
        acc[k] = 0.f;
    for (int t = 0; t < 9; ++t)
    {
        const int q = nbr<false>(x, y, -(t % 3 - 1), -(t / 3 - 1), g, img);
        if (q < 0)
            continue;
        for (int n = 0; n < C; ++n)
        {
            const float gv = gout[(long long)q * C + n];
#pragma unroll
            for (int k = 0; k < C; ++k)
                acc[k] += W[(t * C + n) * C + k] * gv;
        }
    }
#pragma unroll
    for (int k = 0; k < C; ++k)
        gin[p * C + k] = bf2f(hin[p * C + k]) > 0.f ? acc[k] : 0.f;
}

// Weight gradients as FP32 GEMMs over pixels on v_mfma_f32_16x16x4f32
// (A[i][kk]: lane l holds i = l%16, kk = l/16; B[kk][j]: lane l holds
// kk = l/16, j = l%16; D[i][j]: lane l holds i = 4(l/16)+v, j = l%16), one
// MFMA per 4 pixels (kk).  A block of 4 waves takes CH pixels, a quarter
// per wave in pixel order; the waves' sums are added in wave order and the
// block's row goes to its own slot of `part` (k_sum_partials_f32 adds the
// rows in block order), so the gradient is the same on every run.
//
// 16 -> 16 layer: dW[t][n][k] = sum_p gout[p][n] hin[p+off_t][k], one GEMM
// per tap t: A[n][p] = gout[p][n], B[p][k] = hin[p+off_t][k].  Row layout
// [t][n][k].
template <int CH>
__global__ __launch_bounds__(256) void k_bwd16_weight(const float *__restrict__ gout,
                                                      const unsigned short *__restrict__ hin,
                                                      float *__restrict__ part, int stride,
                                                      Geo g)
{
    __shared__ float red[4][9 * C * C];
    const long long npx  = (long long)g.w * g.h * g.nimg;
    const int       lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int       c = lane & 15, pk = lane >> 4;
    v4f             acc[9];
#pragma unroll
    for (int t = 0; t < 9; ++t)
        acc[t] = v4f{ 0.f, 0.f, 0.f, 0.f };
    const long long pb = (long long)blockIdx.x * CH + w * (CH / 4);
    const unsigned  wh_ = (unsigned)g.w * (unsigned)g.h;
    for (int i0 = 0; i0 < CH / 4; i0 += 4)
    {
        const long long p  = pb + i0 + pk;
        const bool      pv = p < npx;
        float           av = 0.f;
        int             x = 0, y = 0, img = 0;
        if (pv)
        {
            av = gout[p * C + c];
            img = (int)((unsigned)p / wh_);
            const int rem = (int)((unsigned)p - (unsigned)img * wh_);
            y = (int)((unsigned)rem / (unsigned)g.w);
            x = rem - y * g.w;
        }
#pragma unroll
        for (int t = 0; t < 9; ++t)
        {
            float bv = 0.f;
            if (pv)
            {
                const int q = nbr<false>(x, y, t % 3 - 1, t / 3 - 1, g, img);
                if (q >= 0)
                    bv = bf2f(hin[(long long)q * C + c]);
            }
            acc[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(av, bv, acc[t], 0, 0, 0);
        }
    }
#pragma unroll
    for (int t = 0; t < 9; ++t)
#pragma unroll
        for (int v = 0; v < 4; ++v)
            red[w][(t * C + 4 * pk + v) * C + c] = acc[t][v];
    __syncthreads();
    float *dst = part + (long long)blockIdx.x * stride;
    for (int i = threadIdx.x; i < 9 * C * C; i += blockDim.x)
        dst[i] = (red[0][i] + red[1][i]) + (red[2][i] + red[3][i]);
}

// out[i] = sum over the nblk partial rows of part[b * stride + i], i < nval,
// in a fixed order: the rows are split into 16 consecutive groups, each
// added in row order by its own thread, and the group sums are added in
// group order (16 values per block)
__global__ __launch_bounds__(256) void k_sum_partials_f32(const float *__restrict__ part,
                                                          int nblk, int stride, int nval,
                                                          float *__restrict__ out)
{
    __shared__ float sh[16][17];
    const int li = threadIdx.x & 15, grp = threadIdx.x >> 4;
    const int i   = blockIdx.x * 16 + li;
    const int per = (nblk + 15) / 16, b0 = grp * per, b1 = min(nblk, b0 + per);
    float     s   = 0.f;
    if (i < nval)
        for (int b = b0; b < b1; ++b)
            s += part[(long long)b * stride + i];
    sh[grp][li] = s;
    __syncthreads();
    if (grp == 0 && i < nval)
    {
        float t = 0.f;
        for (int q = 0; q < 16; ++q)
            t += sh[q][li];
        out[i] = t;
    }
}

// Layers 4 and 1, one GEMM each (tap t = row or column 0..8, 9..15 zero):
//   dW4[t][k] = sum_p g_y[p - off_t] h3[p][k]:  A[t][p] = g_y[p - off_t], B[p][k] = h3[p][k]
//   dW1[t][n] = sum_p g_h1[p][n] x[p + off_t]:  A[n][p] = g_h1[p][n],    B[p][t] = x[p + off_t]
// (the first is sum_p g_y[p] h3[p+off_t][k] with the pixel index shifted).
// Row layout [W4 144 | W1 144].
template <int CH>
__global__ __launch_bounds__(256) void k_bwd_weight_small(const float *__restrict__ gy,
                                                          const unsigned short *__restrict__ h3,
                                                          const float *__restrict__ gh1,
                                                          const unsigned short *__restrict__ x,
                                                          float *__restrict__ part, int stride,
                                                          Geo g)
{
    __shared__ float red[4][2 * C * C];
    const long long npx  = (long long)g.w * g.h * g.nimg;
    const int       lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int       c = lane & 15, pk = lane >> 4;
    const int       tdx = c % 3 - 1, tdy = c / 3 - 1; // tap c (when c < 9)
    v4f             a4 = { 0.f, 0.f, 0.f, 0.f }, a1 = { 0.f, 0.f, 0.f, 0.f };
    const long long pb  = (long long)blockIdx.x * CH + w * (CH / 4);
    const unsigned  wh_ = (unsigned)g.w * (unsigned)g.h;
    for (int i0 = 0; i0 < CH / 4; i0 += 4)
    {
        const long long p  = pb + i0 + pk;
        float           A4 = 0.f, B4 = 0.f, A1 = 0.f, B1 = 0.f;
        if (p < npx)
        {
            const int img = (int)((unsigned)p / wh_);
            const int rem = (int)((unsigned)p - (unsigned)img * wh_);
            const int yy = (int)((unsigned)rem / (unsigned)g.w), xx = rem - yy * g.w;
            if (c < 9)
            {
                const int q4 = nbr<false>(xx, yy, -tdx, -tdy, g, img);
                if (q4 >= 0)
                    A4 = gy[q4];
                const int q1 = nbr<false>(xx, yy, tdx, tdy, g, img);
                if (q1 >= 0)
                    B1 = bf2f(x[q1]);
            }
            B4 = bf2f(h3[p * C + c]);
            A1 = gh1[p * C + c];
        }
        a4 = __builtin_amdgcn_mfma_f32_16x16x4f32(A4, B4, a4, 0, 0, 0);
        a1 = __builtin_amdgcn_mfma_f32_16x16x4f32(A1, B1, a1, 0, 0, 0);
    }
    // lane: D[i = 4 pk + v][j = c]; W4 (t = i, k = c), W1 (t = c, n = i)
#pragma unroll
    for (int v = 0; v < 4; ++v)
    {
        const int i = 4 * pk + v;
        red[w][i * C + c]         = a4[v];
        red[w][C * C + c * C + i] = a1[v];
    }
    __syncthreads();
    float *dst = part + (long long)blockIdx.x * stride;
    for (int i = threadIdx.x; i < NW4 + NW1; i += blockDim.x)
    {
        const int src = i < NW4 ? i : C * C + (i - NW4);
        dst[i]        = (red[0][src] + red[1][src]) + (red[2][src] + red[3][src]);
    }
}

__global__ void k_adam(float *__restrict__ w, const float *__restrict__ gr,
                       float *__restrict__ m, float *__restrict__ v, int n, float lr,
                       float b1, float b2, float c1, float c2, float eps, float clip)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    float gi = gr[i];
    gi       = gi > clip ? clip : (gi < -clip ? -clip : gi);
    m[i]     = b1 * m[i] + (1 - b1) * gi;
    v[i]     = b2 * v[i] + (1 - b2) * gi * gi;
    w[i] -= lr * (m[i] / c1) / (sqrtf(v[i] / c2) + eps);
}

// bf16 copies of W2/W3 in the MFMA B layout [9][16 out][16 in] (same order)
__global__ void k_pack_bf16(const float *__restrict__ w, unsigned short *__restrict__ wb,
                            int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        wb[i] = f2bf(w[i]);
}
// forward uses bf16-rounded weights for layers 1 and 4 as well
__global__ void k_round_bf16(const float *__restrict__ w, float *__restrict__ wr, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        wr[i] = bf2f(f2bf(w[i]));
}

// ---- fused inference: all four layers of one output tile in LDS --------
// Output tile FT x FT, 8 waves.  Each layer is evaluated on the region its
// successor needs (the halo shrinks by one per layer): x on (FT+8)^2, h1 on
// (FT+6)^2, h2 on (FT+4)^2, h3 on (FT+2)^2, the layer-4 partials on
// (FT+2)^2.  Each MFMA tile takes 16 consecutive pixels of the flattened
// region.  Activations sit in LDS as four planes of 4 channels (8 bytes per
// pixel and plane): a lane's B operand (4 channels of one pixel) is one
// 8-byte access, and the 16 lanes that share a channel group touch 128
// consecutive bytes.
// A block's global loads run ahead of their use: the next tile's input halo
// while the current tile's four layers run, the operator diagonal of the
// output pixels during layer 4.
// Zero-padded fields (not PERIODIC) store 0 for every activation outside the
// image, as the layered kernels read their padding.
constexpr int FT     = 28;
constexpr int FX     = FT + 8; // 36
constexpr int FR1    = FT + 6; // 34
constexpr int FR2    = FT + 4; // 32
constexpr int FR3    = FT + 2; // 30
constexpr int NP1    = FR1 * FR1;
constexpr int NP2    = FR2 * FR2;
constexpr int NP3    = FR3 * FR3;
constexpr int F_THR  = 512;
constexpr int F_WAVE = F_THR / 64;
constexpr int F_LDS  = FX * FX * 2 + 4 * NP1 * 8 + 4 * NP2 * 8;
constexpr int F_XL   = (FX * FX + F_THR - 1) / F_THR; // input loads per thread
constexpr int F_OL   = (FT * FT + F_THR - 1) / F_THR; // output pixels per thread
static_assert(4 * NP3 * 8 <= 4 * NP1 * 8, "h3 aliases h1");
static_assert(9 * NP3 * 4 <= 4 * NP2 * 8, "layer-4 partials alias h2");

struct FusedIO
{
    const double *r;    // INFER: residual (FP64)
    const float  *xin;  // raw mode: network input (rounded to bf16 on load)
    const double *part; // INFER: SS_BLOCKS partial sums of r^2 (sigma = rms of r)
    double        n;    // INFER: number of cells of r
    const double *D;    // INFER: operator diagonal
    double       *e;    // INFER: correction
    float        *y;    // raw mode: network output
    double        omega;
};

__device__ inline int wrapi(int v, int n)
{
    v %= n;
    return v < 0 ? v + n : v;
}

// the pixel of the image behind tile-region coordinate (x, y) (already
// offset to image coordinates); -1 outside a zero-padded image
template <bool PERIODIC>
__device__ inline long long region_pixel(int x, int y, const Geo &g, bool big, long long ibase)
{
    if (!PERIODIC)
        return (x < 0 || y < 0 || x >= g.w || y >= g.h) ? -1 : ibase + (long long)y * g.w + x;
    if (big) // the region wraps at most once
    {
        x = x < 0 ? x + g.w : (x >= g.w ? x - g.w : x);
        y = y < 0 ? y + g.h : (y >= g.h ? y - g.h : y);
    }
    else
    {
        x = wrapi(x, g.w);
        y = wrapi(y, g.h);
    }
    return ibase + (long long)y * g.w + x;
}

// One wave's walk over the 16-pixel tiles of a WO x WO region, tiles wave,
// wave + stride, ...: lane m's pixel as region coordinates (i, j), advanced
// without a division per tile.
template <int WO, int STEP>
struct TileWalk
{
    int p, i, j;
    __device__ TileWalk(int first) : p(first), i(first % WO), j(first / WO) {}
    __device__ void next()
    {
        p += STEP;
        i += STEP % WO;
        j += STEP / WO;
        if (i >= WO)
        {
            i -= WO;
            ++j;
        }
    }
};

// a layer's D (lane: channels 4kq..4kq+3 of the tile's pixel at region
// coordinates (i, j), flattened p) with ReLU, one 8-byte store into the four
// planes of the WO x WO region at origin org
template <bool PERIODIC, int WO, int NP>
__device__ inline void store_act(unsigned short *__restrict__ hout, int ox, int oy, int org,
                                 const Geo &g, int p, int i, int j, const v4f &acc, int kq)
{
    if (p >= WO * WO)
        return;
    uint2 h = relu_pack4(acc);
    if (!PERIODIC)
    {
        const int x = ox + org + i, y = oy + org + j;
        if (x < 0 || y < 0 || x >= g.w || y >= g.h)
            h = make_uint2(0u, 0u);
    }
    *reinterpret_cast<uint2 *>(hout + (kq * NP + p) * 4) = h;
}

// A 16 -> 16 layer on the WO x WO region (input (WO+2)^2, row stride WI =
// WO + 2; output row stride WO), one strip of 16 columns x R rows per wave
// and work item.  The wave reads each input row's three column shifts once
// (one 8-byte read per lane: 4 channels of one pixel) and feeds every output
// row that uses it, so an activation is read 3 times instead of 9; each
// output row still receives its 9 taps in tap order.  Columns past WO and
// rows past WO are computed on clamped (in-bounds) reads and not stored.
template <bool PERIODIC, int WO, int NPI, int NPO, int R>
__device__ inline void conv16_rows(const unsigned short *__restrict__ hin,
                                   unsigned short *__restrict__ hout, int ox, int oy,
                                   int org, const Geo &g,
                                   const unsigned short *__restrict__ W, int wave, int m,
                                   int kg)
{
    constexpr int WI = WO + 2, NS = (WO + 15) / 16, NRB = (WO + R - 1) / R;
    v4s           a[9];
#pragma unroll
    for (int t = 0; t < 9; ++t)
        a[t] = w16_tap(W, t, m, kg);
    const unsigned short *src = hin + kg * NPI * 4;
    for (int item = wave; item < NS * NRB; item += F_WAVE)
    {
        const int c  = (item % NS) * 16 + m; // this lane's output column
        const int cc = c < WO ? c : WO - 1;
        const int y0 = (item / NS) * R;
        v4f       acc[R];
#pragma unroll
        for (int y = 0; y < R; ++y)
            acc[y] = v4f{ 0.f, 0.f, 0.f, 0.f };
#pragma unroll
        for (int r = 0; r < R + 2; ++r)
        {
            const int ri = y0 + r < WI ? y0 + r : WI - 1; // input row
#pragma unroll
            for (int dx = 0; dx < 3; ++dx)
            {
                const v4s b = *reinterpret_cast<const v4s *>(src + (ri * WI + cc + dx) * 4);
#pragma unroll
                for (int dy = 0; dy < 3; ++dy)
                    if (r - dy >= 0 && r - dy < R)
                        acc[r - dy] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(
                            a[dy * 3 + dx], b, acc[r - dy], 0, 0, 0);
            }
        }
#pragma unroll
        for (int y = 0; y < R; ++y)
            if (c < WO && y0 + y < WO)
                store_act<PERIODIC, WO, NPO>(hout, ox, oy, org, g, (y0 + y) * WO + c, c, y0 + y,
                                             acc[y], kg);
    }
}

// one thread's input loads for the tile at (ox, oy) of image ibase: the
// input on [-4, FT+4)^2
template <bool PERIODIC, bool INFER>
__device__ inline void fused_loads(const FusedIO &io, const Geo &g, int ox, int oy,
                                   long long ibase, bool big, int tid, double *xv)
{
#pragma unroll
    for (int u = 0; u < F_XL; ++u)
    {
        const int       q = tid + u * F_THR;
        const long long p = q < FX * FX
                                ? region_pixel<PERIODIC>(ox - 4 + q % FX, oy - 4 + q / FX, g, big, ibase)
                                : -1;
        xv[u] = 0.0;
        if (p >= 0)
            xv[u] = INFER ? io.r[p] : (double)io.xin[p];
    }
}

// a thread's output pixel u of the tile at (ox, oy): -1 if outside the image
__device__ inline long long out_pixel(const Geo &g, int ox, int oy, long long ibase, int tid,
                                      int u)
{
    const int q = tid + u * F_THR;
    const int x = ox + q % FT, y = oy + q / FT;
    return (q < FT * FT && x < g.w && y < g.h) ? ibase + (long long)y * g.w + x : -1;
}

// Persistent, two blocks per compute unit (4 waves per SIMD, at most 128
// VGPRs): block b takes tiles b, b + gridDim.x, ... (tile T: column
// T % ntx, row (T / ntx) % nty, image T / (ntx nty)).  The global loads of a
// block's next tile's input are issued right after its current input is in
// LDS, so they are in flight while the four layers of the current tile run.
// LDS reuse across tiles: the next tile writes x and h1 only after every
// wave has passed the current tile's last barrier, and its h2 (which the
// current tile's output phase reads as the partials) only after two more.
template <bool PERIODIC, bool INFER>
__global__ __launch_bounds__(F_THR, 4) void k_net_fused(FusedIO io,
                                                     const unsigned short *__restrict__ Wb,
                                                     Geo g, int ntx, int nty, int ntiles)
{
    __shared__ __align__(16) unsigned char smem[F_LDS];
    unsigned short *sx  = reinterpret_cast<unsigned short *>(smem);
    unsigned short *sh1 = sx + FX * FX;
    unsigned short *sh2 = sh1 + 4 * NP1 * 4;
    unsigned short *sh3 = sh1;                          // h1 is dead once h2 is formed
    float          *sP  = reinterpret_cast<float *>(sh2); // h2 is dead once h3 is formed
    const int       tid = threadIdx.x;
    const bool      big = g.w >= FX && g.h >= FX;
    int             T   = blockIdx.x;
    if (T >= ntiles) // the whole block
        return;
    // sigma from the partial sums: every block adds them in the same order
    double sg = 0.0, inv = 1.0;
    if (INFER)
    {
        double *red = reinterpret_cast<double *>(sh1); // free until the first tile
        double  acc = 0.0;
        for (int i = tid; i < SS_BLOCKS; i += F_THR)
            acc += io.part[i];
        red[tid] = acc;
        __syncthreads();
        for (int off = F_THR / 2; off > 0; off >>= 1)
        {
            if (tid < off)
                red[tid] += red[tid + off];
            __syncthreads();
        }
        sg  = sqrt(red[0] / io.n);
        inv = sg > 0 ? 1.0 / sg : 0.0;
        __syncthreads(); // red[0] is read before the first tile overwrites it
    }
    double xv[F_XL];
    {
        const int img = T / (ntx * nty), rem = T - img * ntx * nty;
        fused_loads<PERIODIC, INFER>(io, g, (rem % ntx) * FT, (rem / ntx) * FT,
                                     (long long)img * g.h * g.w, big, tid, xv);
    }
    for (; T < ntiles; T += gridDim.x)
    {
        const int       img   = T / (ntx * nty), rem = T - img * ntx * nty;
        const int       ox    = (rem % ntx) * FT, oy = (rem / ntx) * FT;
        const long long ibase = (long long)img * g.h * g.w;
#pragma unroll
        for (int u = 0; u < F_XL; ++u)
        {
            const int q = tid + u * F_THR;
            if (q < FX * FX)
                sx[q] = INFER ? f2bf((float)(xv[u] * inv)) : f2bf((float)xv[u]);
        }
        __syncthreads();
        const int Tn = T + (int)gridDim.x;
        if (Tn < ntiles)
        {
            const int imn = Tn / (ntx * nty), rmn = Tn - imn * ntx * nty;
            fused_loads<PERIODIC, INFER>(io, g, (rmn % ntx) * FT, (rmn / ntx) * FT,
                                         (long long)imn * g.h * g.w, big, tid, xv);
        }

        const int lane = tid & 63, wave = tid >> 6, m = lane & 15, kg = lane >> 4;
        // layer 1 on [-3, FT+3)^2: B = the 9 taps of x at the pixel (lane: taps
        // 4kg .. 4kg+3 of pixel m, offsets fixed for the layer)
        {
            const v4s a = w1_operand(Wb + OFF1, m, kg);
            // taps 9..15 (zero weights) read tap 0's neighbour, a pixel past the
            // region's end reads pixel 0: no masked reads (as in the 16 -> 16
            // layers)
            int tof[4];
#pragma unroll
            for (int v = 0; v < 4; ++v)
            {
                const int t = 4 * kg + v;
                tof[v]      = t < 9 ? (t / 3) * FX + t % 3 : 0;
            }
            // two tiles per iteration (independent chains)
            TileWalk<FR1, 32 * F_WAVE> w1(wave * 16 + m), w2((wave + F_WAVE) * 16 + m);
            for (int tile = wave; tile * 16 < FR1 * FR1; tile += 2 * F_WAVE)
            {
                const int base1 = w1.p < FR1 * FR1 ? w1.j * FX + w1.i : 0;
                const int base2 = w2.p < FR1 * FR1 ? w2.j * FX + w2.i : 0;
                v4s       b1, b2;
#pragma unroll
                for (int v = 0; v < 4; ++v)
                {
                    b1[v] = (short)sx[base1 + tof[v]];
                    b2[v] = (short)sx[base2 + tof[v]];
                }
                v4f c1 = { 0.f, 0.f, 0.f, 0.f }, c2 = c1;
                c1 = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b1, c1, 0, 0, 0);
                c2 = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b2, c2, 0, 0, 0);
                store_act<PERIODIC, FR1, NP1>(sh1, ox, oy, -3, g, w1.p, w1.i, w1.j, c1, kg);
                store_act<PERIODIC, FR1, NP1>(sh1, ox, oy, -3, g, w2.p, w2.i, w2.j, c2, kg);
                w1.next();
                w2.next();
            }
        }
        __syncthreads();

        // layers 2 and 3
        conv16_rows<PERIODIC, FR2, NP1, NP2, 4>(sh1, sh2, ox, oy, -2, g, Wb + OFF2, wave, m, kg);
        __syncthreads();
        conv16_rows<PERIODIC, FR3, NP2, NP3, 4>(sh2, sh3, ox, oy, -1, g, Wb + OFF3, wave, m, kg);
        __syncthreads();

        // the operator diagonal at this thread's output pixels, in flight
        // during layer 4's first part
        double dg[F_OL];
#pragma unroll
        for (int u = 0; u < F_OL; ++u)
        {
            const long long op = out_pixel(g, ox, oy, ibase, tid, u);
            dg[u]              = INFER && op >= 0 ? io.D[op] : 1.0;
        }
        // layer 4, part 1: per-tap partials on [-1, FT+1)^2 (lane: taps 4kg ..
        // 4kg+3 of pixel m; rows 9..15 of W4 are zero)
        {
            const v4s a = w4_operand(Wb + OFF4, m, kg);
            for (int tile = wave; tile * 16 < FR3 * FR3; tile += 2 * F_WAVE)
            {
                int pm[2];
                v4s b[2];
                v4f acc[2];
#pragma unroll
                for (int h = 0; h < 2; ++h)
                {
                    pm[h] = (tile + h * F_WAVE) * 16 + m;
                    b[h]  = *reinterpret_cast<const v4s *>(
                        sh3 + (kg * NP3 + (pm[h] < FR3 * FR3 ? pm[h] : 0)) * 4);
                }
#pragma unroll
                for (int h = 0; h < 2; ++h)
                {
                    acc[h] = v4f{ 0.f, 0.f, 0.f, 0.f };
                    acc[h] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b[h], acc[h], 0, 0, 0);
                }
#pragma unroll
                for (int h = 0; h < 2; ++h)
                    if (pm[h] < FR3 * FR3 && kg < 3)
#pragma unroll
                        for (int v = 0; v < 4; ++v)
                            if (4 * kg + v < 9)
                                sP[(4 * kg + v) * NP3 + pm[h]] = acc[h][v];
            }
        }
        __syncthreads();

        // layer 4, part 2: y = sum over taps of the partials at the neighbours
#pragma unroll
        for (int u = 0; u < F_OL; ++u)
        {
            const long long op = out_pixel(g, ox, oy, ibase, tid, u);
            if (op < 0)
                continue;
            const int q = tid + u * F_THR;
            const int i = q % FT, j = q / FT;
            const int x = ox + i, y = oy + j;
            float     s = 0.f;
#pragma unroll
            for (int t = 0; t < 9; ++t)
            {
                const int xx = x + t % 3 - 1, yy = y + t / 3 - 1;
                if (!PERIODIC && (xx < 0 || yy < 0 || xx >= g.w || yy >= g.h))
                    continue;
                s += sP[t * NP3 + (j + t / 3) * FR3 + i + t % 3];
            }
            if (INFER)
                io.e[op] = io.omega * sg * (double)s / dg[u];
            else
                io.y[op] = s;
        }
    }
}

} // namespace

// =========================================================== network state
struct bx_cnn
{
    float          *w      = nullptr; // FP32 master parameters
    float          *wr     = nullptr; // bf16-rounded copy (FP32 storage)
    unsigned short *wb     = nullptr; // bf16 copy (MFMA operands)
    float          *grad   = nullptr;
    float          *adam_m = nullptr, *adam_v = nullptr;
    int             adam_t = 0;
    // operator (for D and the training loss)
    bx_helm *op = nullptr;
    OpRef    opr{};
    double  *D  = nullptr;
    int      onx = 0, ony = 0;
    // inference buffers
    long long       cap = 0, epoch = 0;
    unsigned short *x = nullptr, *h1 = nullptr, *h2 = nullptr, *h3 = nullptr;
    float          *y = nullptr, *P = nullptr; // P: layer-4 partials, 16 per pixel
    double         *part = nullptr; // partial sums of r^2
    // training buffers
    long long tcap = 0;
    float    *gy = nullptr, *g3 = nullptr, *g2 = nullptr, *g1 = nullptr;
    double   *te = nullptr, *tres = nullptr, *tsig = nullptr, *trn2 = nullptr,
           *tloss = nullptr; // [0]: the loss, [1 ..]: per-block partials
    float    *wpart = nullptr; // weight-gradient partials, one row per block
    long long wpcap = 0;
    int      *txy = nullptr;
    int       txycap = 0;
};

static thread_local std::string g_cnn_err;
static int cerr(int code, const std::string &m)
{
    g_cnn_err = m;
    return code;
}
#define C_CHECK(expr)                                                           \
    do                                                                          \
    {                                                                           \
        hipError_t e__ = (expr);                                                \
        if (e__ != hipSuccess)                                                  \
            return cerr(BX_ERR_HIP, std::string(#expr) + ": "                   \
                                        + hipGetErrorString(e__));              \
    } while (0)

extern "C" const char *bx_cnn_last_error(void) { return g_cnn_err.c_str(); }

static void cfree(void *p)
{
    if (p)
        (void)hipFree(p);
}

extern "C" void bx_cnn_destroy(bx_cnn *c)
{
    if (!c)
        return;
    for (void *p : { (void *)c->w, (void *)c->wr, (void *)c->wb, (void *)c->grad,
                     (void *)c->adam_m, (void *)c->adam_v, (void *)c->D,
                     (void *)c->x, (void *)c->h1, (void *)c->h2, (void *)c->h3,
                     (void *)c->y, (void *)c->P, (void *)c->part, (void *)c->gy,
                     (void *)c->g3, (void *)c->g2, (void *)c->g1, (void *)c->te,
                     (void *)c->tres, (void *)c->tsig, (void *)c->trn2,
                     (void *)c->tloss, (void *)c->txy, (void *)c->wpart })
        cfree(p);
    delete c;
}

static int refresh_copies(bx_cnn *c, hipStream_t s)
{
    hipLaunchKernelGGL(k_round_bf16, dim3((NPAR + 255) / 256), dim3(256), 0, s,
                       c->w, c->wr, NPAR);
    hipLaunchKernelGGL(k_pack_bf16, dim3((NPAR + 255) / 256), dim3(256), 0, s,
                       c->w, c->wb, NPAR);
    C_CHECK(hipGetLastError());
    return BX_OK;
}

extern "C" int bx_cnn_create(unsigned seed, bx_cnn **out)
{
    if (!out)
        return cerr(BX_ERR_INVALID, "bx_cnn_create: null");
    bx_cnn *c = new bx_cnn();
    if (hipMalloc((void **)&c->w, NPAR * 4) || hipMalloc((void **)&c->wr, NPAR * 4)
        || hipMalloc((void **)&c->wb, NPAR * 2) || hipMalloc((void **)&c->grad, NPAR * 4)
        || hipMalloc((void **)&c->adam_m, NPAR * 4)
        || hipMalloc((void **)&c->adam_v, NPAR * 4)
        || hipMalloc((void **)&c->part, SS_BLOCKS * 8))
    {
        bx_cnn_destroy(c);
        return cerr(BX_ERR_HIP, "bx_cnn_create: hipMalloc failed");
    }
    // He-style initialisation around a Jacobi-like identity path: the centre
    // tap carries most of the signal so the untrained net is close to
    // e = sigma/D * relu-split identity.
    std::mt19937             rng(seed);
    std::normal_distribution<float> nd(0.f, 1.f);
    std::vector<float>       w(NPAR);
    for (int i = 0; i < NPAR; ++i)
        w[i] = 0.f;
    const float s1 = std::sqrt(2.f / 9.f), s16 = std::sqrt(2.f / (9.f * 16.f));
    for (int i = 0; i < NW1; ++i)
        w[OFF1 + i] = 0.1f * s1 * nd(rng);
    for (int i = 0; i < NW2; ++i)
        w[OFF2 + i] = 0.1f * s16 * nd(rng);
    for (int i = 0; i < NW3; ++i)
        w[OFF3 + i] = 0.1f * s16 * nd(rng);
    for (int i = 0; i < NW4; ++i)
        w[OFF4 + i] = 0.1f * s16 * nd(rng);
    // identity-like path: channel 0 = relu(x), channel 1 = relu(-x)
    w[OFF1 + 4 * C + 0] += 1.f;
    w[OFF1 + 4 * C + 1] -= 1.f;
    for (int L : { OFF2, OFF3 })
    {
        w[L + (4 * C + 0) * C + 0] += 1.f;
        w[L + (4 * C + 1) * C + 1] += 1.f;
    }
    w[OFF4 + 4 * C + 0] += 1.f;
    w[OFF4 + 4 * C + 1] -= 1.f;
    if (hipMemcpy(c->w, w.data(), NPAR * 4, hipMemcpyHostToDevice)
        || hipMemset(c->adam_m, 0, NPAR * 4) || hipMemset(c->adam_v, 0, NPAR * 4))
    {
        bx_cnn_destroy(c);
        return cerr(BX_ERR_HIP, "bx_cnn_create: init copy failed");
    }
    int rc = refresh_copies(c, nullptr);
    if (rc)
    {
        bx_cnn_destroy(c);
        return rc;
    }
    (void)hipDeviceSynchronize();
    *out = c;
    return BX_OK;
}

extern "C" int bx_cnn_set_params(bx_cnn *c, const float *p, int n)
{
    if (!c || !p || n != NPAR)
        return cerr(BX_ERR_INVALID, "bx_cnn_set_params: expected BX_CNN_NPARAMS values");
    C_CHECK(hipMemcpy(c->w, p, NPAR * 4, hipMemcpyHostToDevice));
    C_CHECK(hipMemset(c->adam_m, 0, NPAR * 4));
    C_CHECK(hipMemset(c->adam_v, 0, NPAR * 4));
    c->adam_t = 0;
    int rc    = refresh_copies(c, nullptr);
    C_CHECK(hipDeviceSynchronize());
    return rc;
}

extern "C" int bx_cnn_get_params(bx_cnn *c, float *p, int n)
{
    if (!c || !p || n != NPAR)
        return cerr(BX_ERR_INVALID, "bx_cnn_get_params: expected BX_CNN_NPARAMS values");
    C_CHECK(hipMemcpy(p, c->w, NPAR * 4, hipMemcpyDeviceToHost));
    return BX_OK;
}

// Weight file: 8-byte magic "BXCNN001", uint32 parameter count, uint32
// FNV-1a checksum of the parameter bytes, then the FP32 master parameters
// (little-endian, the order of bx_cnn_get_params).
static uint32_t fnv1a32(const unsigned char *p, size_t n)
{
    uint32_t h = 2166136261u;
    for (size_t i = 0; i < n; ++i)
        h = (h ^ p[i]) * 16777619u;
    return h;
}

extern "C" int bx_cnn_save(bx_cnn *c, const char *path)
{
    if (!c || !path)
        return cerr(BX_ERR_INVALID, "bx_cnn_save: null argument");
    std::vector<float> w(NPAR);
    int                rc;
    if ((rc = bx_cnn_get_params(c, w.data(), NPAR)) != BX_OK)
        return rc;
    const uint32_t n = NPAR;
    const uint32_t h = fnv1a32(reinterpret_cast<const unsigned char *>(w.data()), NPAR * 4);
    FILE          *f = std::fopen(path, "wb");
    if (!f)
        return cerr(BX_ERR_INVALID, std::string("bx_cnn_save: cannot open ") + path);
    const bool ok = std::fwrite("BXCNN001", 1, 8, f) == 8 && std::fwrite(&n, 4, 1, f) == 1
                    && std::fwrite(&h, 4, 1, f) == 1
                    && std::fwrite(w.data(), 4, NPAR, f) == (size_t)NPAR;
    if (std::fclose(f) != 0 || !ok)
        return cerr(BX_ERR_INVALID, std::string("bx_cnn_save: write failed: ") + path);
    return BX_OK;
}

extern "C" int bx_cnn_load(bx_cnn *c, const char *path)
{
    if (!c || !path)
        return cerr(BX_ERR_INVALID, "bx_cnn_load: null argument");
    FILE *f = std::fopen(path, "rb");
    if (!f)
        return cerr(BX_ERR_INVALID, std::string("bx_cnn_load: cannot open ") + path);
    char               magic[8];
    uint32_t           n = 0, h = 0;
    std::vector<float> w(NPAR);
    const bool ok = std::fread(magic, 1, 8, f) == 8 && std::memcmp(magic, "BXCNN001", 8) == 0
                    && std::fread(&n, 4, 1, f) == 1 && n == (uint32_t)NPAR
                    && std::fread(&h, 4, 1, f) == 1
                    && std::fread(w.data(), 4, NPAR, f) == (size_t)NPAR && std::fgetc(f) == EOF;
    std::fclose(f);
    if (!ok)
        return cerr(BX_ERR_INVALID, std::string("bx_cnn_load: not a network weight file of ")
                                       + std::to_string(NPAR) + " parameters: " + path);
    if (fnv1a32(reinterpret_cast<const unsigned char *>(w.data()), NPAR * 4) != h)
        return cerr(BX_ERR_INVALID, std::string("bx_cnn_load: checksum mismatch: ") + path);
    return bx_cnn_set_params(c, w.data(), NPAR);
}

static int ensure_infer(bx_cnn *c, long long npx)
{
    if (npx > 0x7fffffffLL)
        return cerr(BX_ERR_INVALID, "network: more than 2^31 pixels in one launch");
    if (npx <= c->cap)
        return BX_OK;
    for (void *p : { (void *)c->x, (void *)c->h1, (void *)c->h2, (void *)c->h3,
                     (void *)c->y, (void *)c->P })
        cfree(p);
    c->x = c->h1 = c->h2 = c->h3 = nullptr;
    c->y = c->P = nullptr;
    C_CHECK(hipMalloc((void **)&c->x, npx * 2));
    C_CHECK(hipMalloc((void **)&c->h1, npx * C * 2));
    C_CHECK(hipMalloc((void **)&c->h2, npx * C * 2));
    C_CHECK(hipMalloc((void **)&c->h3, npx * C * 2));
    C_CHECK(hipMalloc((void **)&c->y, npx * 4));
    C_CHECK(hipMalloc((void **)&c->P, npx * C * 4));
    c->cap = npx;
    ++c->epoch;
    return BX_OK;
}

int bx_cnn_reserve(bx_cnn *c, long long npx)
{
    if (!c)
        return cerr(BX_ERR_INVALID, "bx_cnn_reserve: null network");
    return ensure_infer(c, npx);
}

long long bx_cnn_epoch(const bx_cnn *c) { return c ? c->epoch : 0; }

// registers the operator: diagonal for the scaling, coefficients for the loss
extern "C" int bx_cnn_set_operator(bx_cnn *c, int nx, int ny, double alpha,
                                   double beta, double dx, double dy,
                                   const double *a, const double *bX,
                                   const double *bY, void *stream)
{
    if (!c || !a || !bX || !bY || nx < 16 || ny < 1)
        return cerr(BX_ERR_INVALID, "bx_cnn_set_operator: bad argument");
    hipStream_t s = (hipStream_t)stream;
    if (c->onx != nx || c->ony != ny)
    {
        cfree(c->D);
        c->D = nullptr;
        C_CHECK(hipMalloc((void **)&c->D, sizeof(double) * nx * ny));
        ++c->epoch;
    }
    c->onx          = nx;
    c->ony          = ny;
    const double dhx = beta / (dx * dx), dhy = beta / (dy * dy);
    hipLaunchKernelGGL(k_diag, dim3((nx + 255) / 256, ny), dim3(256), 0, s, a, bX, bY,
                       c->D, nx, ny, alpha, dhx, dhy);
    C_CHECK(hipGetLastError());
    c->opr = OpRef{ a, bX, bY, c->D, nx, ny, alpha, dhx, dhy };
    return ensure_infer(c, (long long)nx * ny);
}



// The fused network over images g: FT x FT tiles, a persistent grid of two
// blocks per compute unit (the LDS of two blocks fits a CU).
template <bool PERIODIC, bool INFER>
static int launch_fused(const FusedIO &io, const unsigned short *wb, const Geo &g,
                        hipStream_t s)
{
    const int ntx = (g.w + FT - 1) / FT, nty = (g.h + FT - 1) / FT;
    const long long nt = (long long)ntx * nty * g.nimg;
    if (nt > 0x7fffffffLL)
        return cerr(BX_ERR_INVALID, "network: too many tiles in one launch");
    int dev = 0, cus = 0;
    C_CHECK(hipGetDevice(&dev));
    C_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    const int nb = (int)std::min<long long>(nt, 2LL * std::max(cus, 1));
    hipLaunchKernelGGL((k_net_fused<PERIODIC, INFER>), dim3((unsigned)nb), dim3(F_THR), 0, s, io,
                       wb, g, ntx, nty, (int)nt);
    C_CHECK(hipGetLastError());
    return BX_OK;
}

template <bool PER>
static void net_forward(bx_cnn *c, const Geo &g, hipStream_t s)
{
    const long long npx = (long long)g.w * g.h * g.nimg;
    const unsigned  nb  = (unsigned)((npx + 255) / 256);
    const unsigned  nbt = (unsigned)((npx / 16 * 64 + 255) / 256);
    hipLaunchKernelGGL((k_conv1<PER>), dim3(nbt), dim3(256), 0, s, c->x, c->h1,
                       c->wb + OFF1, g);
    hipLaunchKernelGGL((k_conv16<PER>), dim3(nbt), dim3(256), 0, s, c->h1, c->h2,
                       c->wb + OFF2, g);
    hipLaunchKernelGGL((k_conv16<PER>), dim3(nbt), dim3(256), 0, s, c->h2, c->h3,
                       c->wb + OFF3, g);
    hipLaunchKernelGGL(k_conv4_taps, dim3(nbt), dim3(256), 0, s, c->h3, c->P,
                       c->wb + OFF4, npx);
    hipLaunchKernelGGL((k_conv4_sum<PER>), dim3(nb), dim3(256), 0, s, c->P, c->y, g);
}

int bx_cnn_forward(bx_cnn *c, const double *r, double *e, int nx, int ny,
                   double omega, hipStream_t s)
{
    if (!c || !c->D || nx != c->onx || ny != c->ony || nx % 16)
        return cerr(BX_ERR_STATE, "bx_cnn_forward: operator not registered or nx % 16 != 0");
    const long long n = (long long)nx * ny;
    int             rc;
    if ((rc = ensure_infer(c, n)))
        return rc;
    hipLaunchKernelGGL(k_sumsq, dim3(SS_BLOCKS), dim3(256), 0, s, r, n, c->part);
    FusedIO io{ r, nullptr, c->part, (double)n, c->D, e, nullptr, omega };
    return launch_fused<true, true>(io, c->wb, Geo{ nx, ny, 1 }, s);
}

extern "C" int bx_cnn_infer(bx_cnn *c, const double *r, double *e, int nx, int ny,
                            double omega, void *stream)
{
    return bx_cnn_forward(c, r, e, nx, ny, omega, (hipStream_t)stream);
}

// Raw network on a batch of bf16-quantised inputs (zero padding): y = N(x).
// x_in: float host-side values are rounded to bf16 first.  For tests.
extern "C" int bx_cnn_net_forward(bx_cnn *c, const float *x_dev, float *y_dev, int w,
                                  int h, int nimg, int flags, void *stream)
{
    if (!c || !x_dev || !y_dev || w < 1 || h < 1 || nimg < 1)
        return cerr(BX_ERR_INVALID, "bx_cnn_net_forward: bad argument");
    const bool      periodic = (flags & BX_CNN_PERIODIC) != 0;
    const bool      layered  = (flags & BX_CNN_LAYERED) != 0;
    hipStream_t     s = (hipStream_t)stream;
    const long long n = (long long)w * h * nimg;
    const Geo       g{ w, h, nimg };
    int             rc;
    if (!layered)
    {
        if (n > 0x7fffffffLL)
            return cerr(BX_ERR_INVALID, "network: more than 2^31 pixels in one launch");
        FusedIO io{ nullptr, x_dev, nullptr, 0.0, nullptr, nullptr, y_dev, 0.0 };
        if ((rc = periodic ? launch_fused<true, false>(io, c->wb, g, s)
                           : launch_fused<false, false>(io, c->wb, g, s)))
            return rc;
        C_CHECK(hipStreamSynchronize(s));
        return BX_OK;
    }
    if (w % 16)
        return cerr(BX_ERR_INVALID, "bx_cnn_net_forward: layered kernels need w % 16 == 0");
    if ((rc = ensure_infer(c, n)))
        return rc;
    hipLaunchKernelGGL(k_pack_bf16, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s,
                       x_dev, c->x, (int)n);
    if (periodic)
        net_forward<true>(c, g, s);
    else
        net_forward<false>(c, g, s);
    C_CHECK(hipMemcpyAsync(y_dev, c->y, n * 4, hipMemcpyDeviceToDevice, s));
    C_CHECK(hipStreamSynchronize(s));
    return BX_OK;
}

static int ensure_train(bx_cnn *c, long long npx, int batch)
{
    if (npx > c->tcap)
    {
        for (void *p : { (void *)c->gy, (void *)c->g3, (void *)c->g2, (void *)c->g1,
                         (void *)c->te, (void *)c->tres })
            cfree(p);
        C_CHECK(hipMalloc((void **)&c->gy, npx * 4));
        C_CHECK(hipMalloc((void **)&c->g3, npx * C * 4));
        C_CHECK(hipMalloc((void **)&c->g2, npx * C * 4));
        C_CHECK(hipMalloc((void **)&c->g1, npx * C * 4));
        C_CHECK(hipMalloc((void **)&c->te, npx * 8));
        C_CHECK(hipMalloc((void **)&c->tres, npx * 8));
        c->tcap = npx;
    }
    const long long nblk = (npx + WG_CH - 1) / WG_CH;
    if (nblk > c->wpcap)
    {
        cfree(c->wpart);
        c->wpart = nullptr;
        C_CHECK(hipMalloc((void **)&c->wpart, nblk * WG_ROW * 4));
        c->wpcap = nblk;
    }
    if (batch > c->txycap)
    {
        for (void *p : { (void *)c->tsig, (void *)c->trn2, (void *)c->txy, (void *)c->tloss })
            cfree(p);
        C_CHECK(hipMalloc((void **)&c->tsig, batch * 8));
        C_CHECK(hipMalloc((void **)&c->trn2, batch * 8));
        C_CHECK(hipMalloc((void **)&c->txy, batch * 12));
        C_CHECK(hipMalloc((void **)&c->tloss, (1 + LOSS_BLOCKS) * 8));
        c->txycap = batch;
    }
    return ensure_infer(c, npx);
}

// One loss + gradient evaluation on the crops `xy` (device, batch x 3 ints).
// Every sum is formed in a fixed order (per-block partials, then the blocks
// in block order), so the loss and the gradient are the same on every run.
// Loss (summed over the batch, into *loss_dev) and gradient (c->grad) of one
// batch of crops (xy: device, per crop field, x0, y0); no host
// synchronisation.
static int loss_grad(bx_cnn *c, const double *R, int nx, int ny, int batch, int cs,
                     const int *xy, double *loss_dev, hipStream_t s)
{
    const int       border = 5;
    const long long npx    = (long long)batch * cs * cs;
    CropSet         cset{ R, xy, nx, ny, cs };
    Geo             g{ cs, cs, batch };
    const unsigned  nb = (unsigned)((npx + 255) / 256);
    hipLaunchKernelGGL(k_crop_stats, dim3(batch), dim3(256), 0, s, cset, c->tsig,
                       c->trn2, border);
    hipLaunchKernelGGL(k_crop_prep, dim3(2048), dim3(256), 0, s, cset, c->tsig, c->x,
                       batch);
    net_forward<false>(c, g, s);
    hipLaunchKernelGGL(k_crop_e, dim3(2048), dim3(256), 0, s, cset, c->opr, c->tsig,
                       c->y, c->te, batch);
    hipLaunchKernelGGL(k_crop_res, dim3(LOSS_BLOCKS), dim3(256), 0, s, cset, c->opr, c->te,
                       c->trn2, c->tres, c->tloss + 1, batch, border);
    hipLaunchKernelGGL(k_sum_partials_f64, dim3(1), dim3(256), 0, s, c->tloss + 1, LOSS_BLOCKS,
                       loss_dev);
    const float inv_b = 1.0f / (float)batch;
    hipLaunchKernelGGL(k_crop_gy, dim3(2048), dim3(256), 0, s, cset, c->opr, c->tres,
                       c->trn2, c->tsig, c->gy, batch, inv_b);
    const unsigned nblk = (unsigned)((npx + WG_CH - 1) / WG_CH);
    float *const   p3 = c->wpart, *const p2 = c->wpart + NW2, *const p14 = c->wpart + 2 * NW2;
    // layer 4
    hipLaunchKernelGGL(k_bwd4_data, dim3(nb), dim3(256), 0, s, c->gy, c->h3,
                       c->wr + OFF4, c->g3, g);
    // layer 3: g3 is dL/d(h3 pre-activation); weights W3 map h2 -> h3
    hipLaunchKernelGGL(k_bwd16_data, dim3(nb), dim3(256), 0, s, c->g3, c->h2,
                       c->wr + OFF3, c->g2, g);
    hipLaunchKernelGGL((k_bwd16_weight<WG_CH>), dim3(nblk), dim3(256), 0, s, c->g3, c->h2,
                       p3, WG_ROW, g);
    // layer 2: W2 maps h1 -> h2
    hipLaunchKernelGGL(k_bwd16_data, dim3(nb), dim3(256), 0, s, c->g2, c->h1,
                       c->wr + OFF2, c->g1, g);
    hipLaunchKernelGGL((k_bwd16_weight<WG_CH>), dim3(nblk), dim3(256), 0, s, c->g2, c->h1,
                       p2, WG_ROW, g);
    // layers 4 and 1 weights
    hipLaunchKernelGGL((k_bwd_weight_small<WG_CH>), dim3(nblk), dim3(256), 0, s, c->gy, c->h3,
                       c->g1, c->x, p14, WG_ROW, g);
    // the blocks' partials in block order
    hipLaunchKernelGGL(k_sum_partials_f32, dim3((NW2 + 15) / 16), dim3(256), 0, s, p3,
                       (int)nblk, WG_ROW, NW2, c->grad + OFF3);
    hipLaunchKernelGGL(k_sum_partials_f32, dim3((NW2 + 15) / 16), dim3(256), 0, s, p2,
                       (int)nblk, WG_ROW, NW2, c->grad + OFF2);
    hipLaunchKernelGGL(k_sum_partials_f32, dim3((NW4 + 15) / 16), dim3(256), 0, s, p14, (int)nblk, WG_ROW,
                       NW4, c->grad + OFF4);
    hipLaunchKernelGGL(k_sum_partials_f32, dim3((NW1 + 15) / 16), dim3(256), 0, s, p14 + NW4, (int)nblk,
                       WG_ROW, NW1, c->grad + OFF1);
    C_CHECK(hipGetLastError());
    return BX_OK;
}

extern "C" int bx_cnn_loss_grad(bx_cnn *c, const double *R, int nfields, int nx,
                                int ny, const int *crops_host, int batch, int crop,
                                double *loss, float *grad_host, void *stream)
{
    if (!c || !c->D || !R || !crops_host || batch < 1 || crop < 16 || crop % 16
        || !loss)
        return cerr(BX_ERR_INVALID, "bx_cnn_loss_grad: bad argument");
    if (nx != c->onx || ny != c->ony)
        return cerr(BX_ERR_STATE, "bx_cnn_loss_grad: field size differs from the operator");
    for (int b = 0; b < batch; ++b)
        if (crops_host[3 * b] < 0 || crops_host[3 * b] >= nfields)
            return cerr(BX_ERR_INVALID, "bx_cnn_loss_grad: crop field index out of range");
    hipStream_t s = (hipStream_t)stream;
    int         rc;
    if ((rc = ensure_train(c, (long long)batch * crop * crop, batch)))
        return rc;
    C_CHECK(hipMemcpyAsync(c->txy, crops_host, batch * 12, hipMemcpyHostToDevice, s));
    if ((rc = loss_grad(c, R, nx, ny, batch, crop, c->txy, c->tloss, s)))
        return rc;
    double l = 0;
    C_CHECK(hipMemcpyAsync(&l, c->tloss, 8, hipMemcpyDeviceToHost, s));
    C_CHECK(hipStreamSynchronize(s));
    *loss = l / batch;
    if (grad_host)
    {
        C_CHECK(hipMemcpy(grad_host, c->grad, NPAR * 4, hipMemcpyDeviceToHost));
    }
    return BX_OK;
}

static int train_steps(bx_cnn *c, const double *R, int nx, int ny, int crop, int batch,
                       int steps, float lr, const std::vector<int> &xy, int *dxy,
                       double *dloss, double *loss_history, hipStream_t s);

extern "C" int bx_cnn_train(bx_cnn *c, const double *R, int nfields, int nx, int ny,
                            int crop, int batch, int steps, float lr, unsigned seed,
                            double *loss_history, void *stream)
{
    if (!c || !c->D || !R || nfields < 1 || batch < 1 || steps < 1 || crop < 16
        || crop % 16 || !(lr > 0))
        return cerr(BX_ERR_INVALID, "bx_cnn_train: bad argument");
    if (nx != c->onx || ny != c->ony)
        return cerr(BX_ERR_STATE, "bx_cnn_train: field size differs from the operator");
    hipStream_t s = (hipStream_t)stream;
    int         rc;
    if ((rc = ensure_train(c, (long long)batch * crop * crop, batch)))
        return rc;
    // every step's crops drawn up front and uploaded once, the losses kept
    // on the device: the steps run without host synchronisation
    std::mt19937     rng(seed);
    std::vector<int> xy(3LL * batch * steps);
    for (long long i = 0; i < (long long)batch * steps; ++i)
    {
        xy[3 * i]     = (int)(rng() % (unsigned)nfields);
        xy[3 * i + 1] = (int)(rng() % (unsigned)nx);
        xy[3 * i + 2] = (int)(rng() % (unsigned)ny);
    }
    int    *dxy   = nullptr;
    double *dloss = nullptr;
    if (hipMalloc((void **)&dxy, xy.size() * sizeof(int)) != hipSuccess
        || hipMalloc((void **)&dloss, (size_t)steps * sizeof(double)) != hipSuccess)
    {
        cfree(dxy);
        return cerr(BX_ERR_HIP, "bx_cnn_train: hipMalloc failed");
    }
    rc = train_steps(c, R, nx, ny, crop, batch, steps, lr, xy, dxy, dloss, loss_history, s);
    cfree(dxy);
    cfree(dloss);
    return rc;
}

static int train_steps(bx_cnn *c, const double *R, int nx, int ny, int crop, int batch,
                       int steps, float lr, const std::vector<int> &xy, int *dxy,
                       double *dloss, double *loss_history, hipStream_t s)
{
    int         rc;
    const float b1 = 0.9f, b2 = 0.999f;
    C_CHECK(hipMemcpyAsync(dxy, xy.data(), xy.size() * sizeof(int), hipMemcpyHostToDevice, s));
    for (int st = 0; st < steps; ++st)
    {
        if ((rc = loss_grad(c, R, nx, ny, batch, crop, dxy + 3LL * batch * st, dloss + st, s)))
            return rc;
        ++c->adam_t;
        const float c1 = 1.f - std::pow(b1, (float)c->adam_t);
        const float c2 = 1.f - std::pow(b2, (float)c->adam_t);
        hipLaunchKernelGGL(k_adam, dim3((NPAR + 255) / 256), dim3(256), 0, s, c->w,
                           c->grad, c->adam_m, c->adam_v, NPAR, lr, b1, b2, c1, c2,
                           1e-8f, 10.f);
        if ((rc = refresh_copies(c, s)))
            return rc;
    }
    std::vector<double> l(steps);
    C_CHECK(hipMemcpyAsync(l.data(), dloss, (size_t)steps * sizeof(double), hipMemcpyDeviceToHost,
                           s));
    C_CHECK(hipStreamSynchronize(s));
    if (loss_history)
        for (int st = 0; st < steps; ++st)
            loss_history[st] = l[st] / batch;
    return BX_OK;
}
