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
// All four layers run on v_mfma_f32_16x16x16bf16_1k (see the forward
// kernels): one wave takes 16 pixels (M); the 16 -> 16 layers accumulate 9
// MFMAs (one per tap, K = 16 input channels), the first layer takes the 9
// taps as K, the last one forms 9 per-tap partials per pixel as N.
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

#include <cmath>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "../../include/bamrex_hip.h"
#include "bx_helmholtz.h"

namespace {

typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

constexpr int C = 16;
constexpr int NW1 = 9 * C, NW2 = 9 * C * C, NW3 = 9 * C * C, NW4 = 9 * C;
constexpr int OFF1 = 0, OFF2 = NW1, OFF3 = OFF2 + NW2, OFF4 = OFF3 + NW3;
constexpr int NPAR = OFF4 + NW4;
static_assert(NPAR == BX_CNN_NPARAMS, "parameter count");
// training reductions: pixels per weight-gradient block, the row of one
// block's partials (W3 | W2 | W4 | W1), blocks of the loss sum
constexpr int WG_CH       = 128;
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
// Every layer runs on v_mfma_f32_16x16x16bf16_1k with one wave per 16 pixels:
//   A[i][k] (lane l: i = l%16, k = 4(l/16)+v), B[k][j] (lane l: j = l%16),
//   D[i][j] (lane l: i = 4(l/16)+v, j = l%16).
// Each D[i][j] is a dot product over k of row i of A with column j of B, so a
// pixel's result does not depend on which pixels share its tile: the fused
// inference kernel below and these layer-by-layer kernels agree bit for bit.
//   layer 1: A = the 9 taps of x at a pixel (k = tap, 7 zero), B = W1 (tap x out)
//   layers 2, 3: A = 16 channels of a neighbour, B = W (in x out), 9 MFMAs
//   layer 4: A = 16 channels of h3 at a pixel, B = W4 (in x tap), giving the
//            per-tap partials P[pixel][tap]; y = sum over taps of P at the
//            neighbours, in tap order.

// B operand of layer 1: B[k = tap][j = out] = W1[tap][out] (tap < 9)
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
// B operand of layer 4: B[k = in channel][j = tap] = W4[tap][in] (tap < 9)
__device__ inline v4s w4_operand(const unsigned short *__restrict__ Wb4, int m, int kg)
{
    v4s b;
#pragma unroll
    for (int v = 0; v < 4; ++v)
        b[v] = m < 9 ? (short)Wb4[m * C + 4 * kg + v] : (short)0;
    return b;
}

// layer 1: x (bf16, 1 channel) -> h1 (bf16, 16, ReLU)
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
    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, w1_operand(Wb1, m, kg), acc, 0, 0, 0);
#pragma unroll
    for (int v = 0; v < 4; ++v)
    {
        const float s = acc[v];
        hout[(p0 + 4 * kg + v) * C + m] = f2bf(s > 0.f ? s : 0.f);
    }
}

// layers 2, 3: h (bf16, 16) -> h' (bf16, 16, ReLU) on MFMA.
// Wb: bf16 weights [9][16 out][16 in].
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
        v4s       a = { 0, 0, 0, 0 };
        if (q >= 0)
            a = *reinterpret_cast<const v4s *>(hin + (long long)q * C + 4 * kg);
        const v4s b
            = *reinterpret_cast<const v4s *>(Wb + (t * C + m) * C + 4 * kg);
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
    }
    // lane holds D[4*kg + v][m]: pixel p0 + 4*kg + v, channel m
#pragma unroll
    for (int v = 0; v < 4; ++v)
    {
        const float s = acc[v];
        hout[(p0 + 4 * kg + v) * C + m] = f2bf(s > 0.f ? s : 0.f);
    }
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
    const v4s a = *reinterpret_cast<const v4s *>(hin + (p0 + m) * C + 4 * kg);
    v4f       acc = { 0.f, 0.f, 0.f, 0.f };
    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, w4_operand(Wb4, m, kg), acc, 0, 0, 0);
#pragma unroll
    for (int v = 0; v < 4; ++v)
        P[(p0 + 4 * kg + v) * C + m] = acc[v];
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

// x = bf16(r * inv_sigma)
__global__ __launch_bounds__(256) void k_prep_full(const double *__restrict__ r,
                                                   unsigned short *__restrict__ x,
                                                   long long n,
                                                   const double *__restrict__ ss)
{
    const double inv = ss[0] > 0 ? 1.0 / ss[0] : 0.0;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        x[i] = f2bf((float)(r[i] * inv));
}

// e = omega * sigma * y / D
__global__ __launch_bounds__(256) void k_finish_full(const float *__restrict__ y,
                                                     const double *__restrict__ Dg,
                                                     double *__restrict__ e,
                                                     long long n, double omega,
                                                     const double *__restrict__ ss)
{
    const double sg = ss[0];
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        e[i] = omega * sg * (double)y[i] / Dg[i];
}

// rms of r (two-stage)
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
__global__ void k_rms_final(const double *__restrict__ partial, int nb,
                            long long n, double *__restrict__ out)
{
    __shared__ double sh[256];
    double            s = 0;
    for (int i = threadIdx.x; i < nb; i += blockDim.x)
        s += partial[i];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1)
    {
        if ((int)threadIdx.x < off)
            sh[threadIdx.x] += sh[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        out[0] = sqrt(sh[0] / (double)n);
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

// out = sum of n partials, one thread, in index order
__global__ void k_sum_partials_f64(const double *__restrict__ part, int n,
                                   double *__restrict__ out)
{
    if (threadIdx.x == 0 && blockIdx.x == 0)
    {
        double s = 0;
        for (int i = 0; i < n; ++i)
            s += part[i];
        *out = s;
    }
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

// weight gradient of a 16->16 layer: dW[t][n][k] = sum_p gout[p][n] hin[p+off_t][k]
// block: 256 threads, thread = (n, k), 9 taps each; chunk of CH pixels, in
// pixel order; the block's sums go to its own slot of `part` (layout
// [block][t][n][k]) and k_sum_partials_f32 adds the blocks in block order,
// so the gradient is the same on every run
template <int CH>
__global__ __launch_bounds__(256) void k_bwd16_weight(const float *__restrict__ gout,
                                                      const unsigned short *__restrict__ hin,
                                                      float *__restrict__ part, int stride,
                                                      Geo g)
{
    const long long npx = (long long)g.w * g.h * g.nimg;
    const int       n = threadIdx.x >> 4, k = threadIdx.x & 15;
    float           acc[9];
#pragma unroll
    for (int t = 0; t < 9; ++t)
        acc[t] = 0.f;
    const long long pbeg = (long long)blockIdx.x * CH;
    for (long long p = pbeg; p < pbeg + CH && p < npx; ++p)
    {
        const float gv = gout[p * C + n];
        if (gv == 0.f)
            continue;
        const unsigned wh_ = (unsigned)g.w * (unsigned)g.h;
        const int img = (int)((unsigned)p / wh_);
        const int rem = (int)((unsigned)p - (unsigned)img * wh_);
        const int y = (int)((unsigned)rem / (unsigned)g.w), x = rem - y * g.w;
#pragma unroll
        for (int t = 0; t < 9; ++t)
        {
            const int q = nbr<false>(x, y, t % 3 - 1, t / 3 - 1, g, img);
            if (q >= 0)
                acc[t] += gv * bf2f(hin[(long long)q * C + k]);
        }
    }
    float *dst = part + (long long)blockIdx.x * stride;
#pragma unroll
    for (int t = 0; t < 9; ++t)
        dst[(t * C + n) * C + k] = acc[t];
}

// out[i] = sum over the nblk partial rows of part[b * stride + i], i < nval,
// in block order
__global__ __launch_bounds__(256) void k_sum_partials_f32(const float *__restrict__ part,
                                                          int nblk, int stride, int nval,
                                                          float *__restrict__ out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nval)
        return;
    float s = 0.f;
    for (int b = 0; b < nblk; ++b)
        s += part[(long long)b * stride + i];
    out[i] = s;
}

// weight gradient of layer 4 (dW4[t][k] += sum_p g_y[p] h3[p+off_t][k]) and
// layer 1 (dW1[t][n] += sum_p g_h1[p][n] x[p+off_t]); thread = (t, c)
// (per-block partials, layout [block][W4 144 | W1 144], as for the 16->16 layers)
template <int CH>
__global__ __launch_bounds__(320) void k_bwd_weight_small(const float *__restrict__ gy,
                                                          const unsigned short *__restrict__ h3,
                                                          const float *__restrict__ gh1,
                                                          const unsigned short *__restrict__ x,
                                                          float *__restrict__ part, int stride,
                                                          Geo g)
{
    const long long npx = (long long)g.w * g.h * g.nimg;
    const int       tid = threadIdx.x; // 0..143 for W4, 144 threads for W1 (second block half)
    const bool      w4  = tid < 144;
    const int       idx = w4 ? tid : tid - 144;
    if (idx >= 144)
        return;
    const int t = idx / C, c = idx % C;
    float     acc = 0.f;
    const long long pbeg = (long long)blockIdx.x * CH;
    for (long long p = pbeg; p < pbeg + CH && p < npx; ++p)
    {
        const unsigned wh_ = (unsigned)g.w * (unsigned)g.h;
        const int img = (int)((unsigned)p / wh_);
        const int rem = (int)((unsigned)p - (unsigned)img * wh_);
        const int y = (int)((unsigned)rem / (unsigned)g.w), xx = rem - y * g.w;
        const int q = nbr<false>(xx, y, t % 3 - 1, t / 3 - 1, g, img);
        if (q < 0)
            continue;
        if (w4)
            acc += gy[p] * bf2f(h3[(long long)q * C + c]);
        else
            acc += gh1[p * C + c] * bf2f(x[q]);
    }
    part[(long long)blockIdx.x * stride + tid] = acc;
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
// Output tile FT x FT.  Each layer is evaluated on the region its successor
// needs (the halo shrinks by one per layer): x on (FT+8)^2, h1 on (FT+6)^2,
// h2 on (FT+4)^2, h3 on (FT+2)^2, the layer-4 partials on (FT+2)^2.  Each
// MFMA tile takes 16 consecutive pixels of the flattened region.  Activations
// sit in LDS as four planes of 4 channels (8 bytes per pixel and plane), with
// plane strides of 16 (mod 32) pixels, so the 8-byte A-operand reads of the
// 64 lanes fall in distinct banks.  Zero-padded fields (not PERIODIC) store 0
// for every activation outside the image, as the layered kernels read their
// padding.
constexpr int FT   = 28;
constexpr int FX   = FT + 8; // 36
constexpr int FR1  = FT + 6; // 34
constexpr int FR2  = FT + 4; // 32
constexpr int FR3  = FT + 2; // 30
constexpr int NP1  = 1168;   // >= FR1^2, = 16 (mod 32)
constexpr int NP2  = 1040;   // >= FR2^2
constexpr int NP3  = 912;    // >= FR3^2
static_assert(NP1 >= FR1 * FR1 && NP2 >= FR2 * FR2 && NP3 >= FR3 * FR3, "planes");
constexpr int F_LDS = FX * FX * 2 + 4 * NP1 * 8 + 4 * NP2 * 8;
static_assert(4 * NP3 * 8 <= 4 * NP1 * 8, "h3 aliases h1");
static_assert(9 * NP3 * 4 <= 4 * NP2 * 8, "layer-4 partials alias h2");

struct FusedIO
{
    const double *r;   // INFER: residual (FP64)
    const float  *xin; // raw mode: network input (rounded to bf16 on load)
    const double *ss;  // INFER: sigma (rms of r)
    const double *D;   // INFER: operator diagonal
    double       *e;   // INFER: correction
    float        *y;   // raw mode: network output
    double        omega;
};

__device__ inline int wrapi(int v, int n)
{
    v %= n;
    return v < 0 ? v + n : v;
}

// D of a 16 -> 16 or 1 -> 16 layer (lane: pixels 4kg+v of the tile,
// channel m) with ReLU, into the planes of the WO x WO region at origin org
template <bool PERIODIC, int WO, int NP>
__device__ inline void store_act(unsigned short *__restrict__ hout, int ox, int oy, int org,
                                 const Geo &g, int tile, const v4f &acc, int m, int kg)
{
#pragma unroll
    for (int v = 0; v < 4; ++v)
    {
        const int po = tile * 16 + 4 * kg + v;
        if (po >= WO * WO)
            continue;
        const float    sv = acc[v];
        unsigned short h  = f2bf(sv > 0.f ? sv : 0.f);
        if (!PERIODIC)
        {
            const int x = ox + org + po % WO, y = oy + org + po / WO;
            if (x < 0 || y < 0 || x >= g.w || y >= g.h)
                h = 0;
        }
        hout[((m >> 2) * NP + po) * 4 + (m & 3)] = h;
    }
}

// A 16 -> 16 layer on the WO x WO region (input (WO+2)^2): 9 MFMAs (one per
// tap) per 16 output pixels, two tiles interleaved per wave.
template <bool PERIODIC, int WO, int NPI, int NPO>
__device__ inline void conv16_tiles(const unsigned short *__restrict__ hin,
                                    unsigned short *__restrict__ hout, int ox, int oy,
                                    int org, const Geo &g,
                                    const unsigned short *__restrict__ W, int wave, int m,
                                    int kg)
{
    constexpr int WI = WO + 2, NPX = WO * WO, NT = (NPX + 15) / 16;
    v4s           b[9];
#pragma unroll
    for (int t = 0; t < 9; ++t)
        b[t] = *reinterpret_cast<const v4s *>(W + (t * C + m) * C + 4 * kg);
    for (int tile = wave; tile < NT; tile += 8)
    {
        const int  t2  = tile + 4;
        const int  p1  = tile * 16 + m, p2 = t2 * 16 + m;
        const bool a1v = p1 < NPX, a2v = p2 < NPX;
        const int  o1  = a1v ? (p1 / WO) * WI + p1 % WO : 0;
        const int  o2  = a2v ? (p2 / WO) * WI + p2 % WO : 0;
        v4f        c1 = { 0.f, 0.f, 0.f, 0.f }, c2 = { 0.f, 0.f, 0.f, 0.f };
#pragma unroll
        for (int t = 0; t < 9; ++t)
        {
            const int d  = (t / 3) * WI + t % 3;
            v4s       a1 = { 0, 0, 0, 0 }, a2 = { 0, 0, 0, 0 };
            if (a1v)
                a1 = *reinterpret_cast<const v4s *>(hin + (kg * NPI + o1 + d) * 4);
            if (a2v)
                a2 = *reinterpret_cast<const v4s *>(hin + (kg * NPI + o2 + d) * 4);
            c1 = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a1, b[t], c1, 0, 0, 0);
            c2 = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a2, b[t], c2, 0, 0, 0);
        }
        store_act<PERIODIC, WO, NPO>(hout, ox, oy, org, g, tile, c1, m, kg);
        if (t2 < NT)
            store_act<PERIODIC, WO, NPO>(hout, ox, oy, org, g, t2, c2, m, kg);
    }
}

template <bool PERIODIC, bool INFER>
__global__ __launch_bounds__(256) void k_net_fused(FusedIO io,
                                                   const unsigned short *__restrict__ Wb, Geo g)
{
    __shared__ __align__(16) unsigned char smem[F_LDS];
    unsigned short *sx  = reinterpret_cast<unsigned short *>(smem);
    unsigned short *sh1 = sx + FX * FX;
    unsigned short *sh2 = sh1 + 4 * NP1 * 4;
    unsigned short *sh3 = sh1;                          // h1 is dead once h2 is formed
    float          *sP  = reinterpret_cast<float *>(sh2); // h2 is dead once h3 is formed
    const int ox = blockIdx.x * FT, oy = blockIdx.y * FT, img = blockIdx.z;
    const int tid = threadIdx.x;
    const long long ibase = (long long)img * g.h * g.w;
    auto inside = [&](int x, int y) { return x >= 0 && y >= 0 && x < g.w && y < g.h; };

    // x = bf16(r / sigma) on [-4, FT+4)^2
    double inv = 1.0;
    if (INFER)
        inv = io.ss[0] > 0 ? 1.0 / io.ss[0] : 0.0;
    const bool big = g.w >= FX && g.h >= FX;
    for (int q = tid; q < FX * FX; q += blockDim.x)
    {
        const int i = q % FX, j = q / FX;
        int       x = ox - 4 + i, y = oy - 4 + j;
        unsigned short v = 0;
        if (PERIODIC || inside(x, y))
        {
            if (PERIODIC)
            {
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
            }
            const long long p = ibase + (long long)y * g.w + x;
            v = INFER ? f2bf((float)(io.r[p] * inv)) : f2bf(io.xin[p]);
        }
        sx[q] = v;
    }
    __syncthreads();

    const int lane = tid & 63, wave = tid >> 6, m = lane & 15, kg = lane >> 4;
    // layer 1 on [-3, FT+3)^2: A = the 9 taps of x at the pixel
    {
        const v4s b = w1_operand(Wb + OFF1, m, kg);
        for (int tile = wave; tile * 16 < FR1 * FR1; tile += 4)
        {
            const int pm = tile * 16 + m;
            v4s       a  = { 0, 0, 0, 0 };
            if (pm < FR1 * FR1)
            {
                const int im = pm % FR1, jm = pm / FR1;
#pragma unroll
                for (int v = 0; v < 4; ++v)
                {
                    const int t = 4 * kg + v;
                    a[v] = t < 9 ? (short)sx[(jm + t / 3) * FX + im + t % 3] : (short)0;
                }
            }
            v4f acc = { 0.f, 0.f, 0.f, 0.f };
            acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
            store_act<PERIODIC, FR1, NP1>(sh1, ox, oy, -3, g, tile, acc, m, kg);
        }
    }
    __syncthreads();

    // layers 2 and 3
    conv16_tiles<PERIODIC, FR2, NP1, NP2>(sh1, sh2, ox, oy, -2, g, Wb + OFF2, wave, m, kg);
    __syncthreads();
    conv16_tiles<PERIODIC, FR3, NP2, NP3>(sh2, sh3, ox, oy, -1, g, Wb + OFF3, wave, m, kg);
    __syncthreads();

    // layer 4, part 1: per-tap partials on [-1, FT+1)^2
    {
        const v4s b = w4_operand(Wb + OFF4, m, kg);
        for (int tile = wave; tile * 16 < FR3 * FR3; tile += 4)
        {
            const int pm = tile * 16 + m;
            v4s       a  = { 0, 0, 0, 0 };
            if (pm < FR3 * FR3)
                a = *reinterpret_cast<const v4s *>(sh3 + (kg * NP3 + pm) * 4);
            v4f acc = { 0.f, 0.f, 0.f, 0.f };
            acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
            if (m < 9)
#pragma unroll
                for (int v = 0; v < 4; ++v)
                {
                    const int po = tile * 16 + 4 * kg + v;
                    if (po < FR3 * FR3)
                        sP[m * NP3 + po] = acc[v];
                }
        }
    }
    __syncthreads();

    // layer 4, part 2: y = sum over taps of the partials at the neighbours
    for (int q = tid; q < FT * FT; q += blockDim.x)
    {
        const int i = q % FT, j = q / FT;
        const int x = ox + i, y = oy + j;
        if (x >= g.w || y >= g.h)
            continue;
        float s = 0.f;
#pragma unroll
        for (int t = 0; t < 9; ++t)
        {
            if (!PERIODIC && !inside(x + t % 3 - 1, y + t / 3 - 1))
                continue;
            s += sP[t * NP3 + (j + t / 3) * FR3 + i + t % 3];
        }
        const long long p = ibase + (long long)y * g.w + x;
        if (INFER)
            io.e[p] = io.omega * io.ss[0] * (double)s / io.D[p];
        else
            io.y[p] = s;
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
    double         *sig = nullptr, *part = nullptr;
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
                     (void *)c->y, (void *)c->P, (void *)c->sig, (void *)c->part, (void *)c->gy,
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
        || hipMalloc((void **)&c->sig, 64) || hipMalloc((void **)&c->part, 1024 * 8))
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



static dim3 fused_grid(const Geo &g)
{
    return dim3((unsigned)((g.w + FT - 1) / FT), (unsigned)((g.h + FT - 1) / FT),
                (unsigned)g.nimg);
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
    hipLaunchKernelGGL(k_sumsq, dim3(512), dim3(256), 0, s, r, n, c->part);
    hipLaunchKernelGGL(k_rms_final, dim3(1), dim3(256), 0, s, c->part, 512, n, c->sig);
    FusedIO io{ r, nullptr, c->sig, c->D, e, nullptr, omega };
    hipLaunchKernelGGL((k_net_fused<true, true>), fused_grid(Geo{ nx, ny, 1 }), dim3(256),
                       0, s, io, c->wb, Geo{ nx, ny, 1 });
    C_CHECK(hipGetLastError());
    return BX_OK;
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
        FusedIO io{ nullptr, x_dev, nullptr, nullptr, nullptr, y_dev, 0.0 };
        if (periodic)
            hipLaunchKernelGGL((k_net_fused<true, false>), fused_grid(g), dim3(256), 0, s,
                               io, c->wb, g);
        else
            hipLaunchKernelGGL((k_net_fused<false, false>), fused_grid(g), dim3(256), 0, s,
                               io, c->wb, g);
        C_CHECK(hipGetLastError());
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
static int loss_grad(bx_cnn *c, const double *R, int nfields, int nx, int ny,
                     int batch, int cs, double *loss_host, hipStream_t s)
{
    (void)nfields;
    const int       border = 5;
    const long long npx    = (long long)batch * cs * cs;
    CropSet         cset{ R, c->txy, nx, ny, cs };
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
    hipLaunchKernelGGL(k_sum_partials_f64, dim3(1), dim3(64), 0, s, c->tloss + 1, LOSS_BLOCKS,
                       c->tloss);
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
    hipLaunchKernelGGL((k_bwd_weight_small<WG_CH>), dim3(nblk), dim3(288), 0, s, c->gy, c->h3,
                       c->g1, c->x, p14, WG_ROW, g);
    // the blocks' partials in block order
    hipLaunchKernelGGL(k_sum_partials_f32, dim3((NW2 + 255) / 256), dim3(256), 0, s, p3,
                       (int)nblk, WG_ROW, NW2, c->grad + OFF3);
    hipLaunchKernelGGL(k_sum_partials_f32, dim3((NW2 + 255) / 256), dim3(256), 0, s, p2,
                       (int)nblk, WG_ROW, NW2, c->grad + OFF2);
    hipLaunchKernelGGL(k_sum_partials_f32, dim3(1), dim3(256), 0, s, p14, (int)nblk, WG_ROW,
                       NW4, c->grad + OFF4);
    hipLaunchKernelGGL(k_sum_partials_f32, dim3(1), dim3(256), 0, s, p14 + NW4, (int)nblk,
                       WG_ROW, NW1, c->grad + OFF1);
    C_CHECK(hipGetLastError());
    double l = 0;
    C_CHECK(hipMemcpyAsync(&l, c->tloss, 8, hipMemcpyDeviceToHost, s));
    C_CHECK(hipStreamSynchronize(s));
    *loss_host = l / batch;
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
    if ((rc = loss_grad(c, R, nfields, nx, ny, batch, crop, loss, s)))
        return rc;
    if (grad_host)
    {
        C_CHECK(hipMemcpy(grad_host, c->grad, NPAR * 4, hipMemcpyDeviceToHost));
    }
    return BX_OK;
}

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
    std::mt19937     rng(seed);
    std::vector<int> xy(3 * batch);
    const float      b1 = 0.9f, b2 = 0.999f;
    for (int st = 0; st < steps; ++st)
    {
        for (int b = 0; b < batch; ++b)
        {
            xy[3 * b]     = (int)(rng() % (unsigned)nfields);
            xy[3 * b + 1] = (int)(rng() % (unsigned)nx);
            xy[3 * b + 2] = (int)(rng() % (unsigned)ny);
        }
        C_CHECK(hipMemcpyAsync(c->txy, xy.data(), batch * 12, hipMemcpyHostToDevice, s));
        double l = 0;
        if ((rc = loss_grad(c, R, nfields, nx, ny, batch, crop, &l, s)))
            return rc;
        if (loss_history)
            loss_history[st] = l;
        ++c->adam_t;
        const float c1 = 1.f - std::pow(b1, (float)c->adam_t);
        const float c2 = 1.f - std::pow(b2, (float)c->adam_t);
        hipLaunchKernelGGL(k_adam, dim3((NPAR + 255) / 256), dim3(256), 0, s, c->w,
                           c->grad, c->adam_m, c->adam_v, NPAR, lr, b1, b2, c1, c2,
                           1e-8f, 10.f);
        if ((rc = refresh_copies(c, s)))
            return rc;
    }
    C_CHECK(hipStreamSynchronize(s));
    return BX_OK;
}
