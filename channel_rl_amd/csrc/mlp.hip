// mlp.hip -- the learned wall-stress closure: a 4 -> 64 -> 64 -> 64 -> 2
// multilayer perceptron with SiLU activations, BF16 weights and FP32
// accumulation, evaluated and trained on the GPU with hand-written kernels.
// The hidden 64x64 layers (forward, the input-gradient products of the
// backward pass, and the weight gradients) run on the matrix cores
// (v_mfma_f32_16x16x32_bf16); the 4-wide input layer and 2-wide output layer
// are plain FP32 arithmetic.  No reference counterpart: the reference's wall
// stress comes from the host (alpha_x, code/smarties_stat.f90:231-258).
//
// Operand layout of v_mfma_f32_16x16x32_bf16 (one wave, lane l, g = l/16):
//   A (16 x 32): lane holds A[l%16][8g + e], e = 0..7
//   B (32 x 16): lane holds B[8g + e][l%16]
//   D (16 x 16): lane holds D[4g + r][l%16], r = 0..3
// A layer is computed feature-major, Z^T = W^T H^T, with the 16 samples of a
// group as the columns, so a lane holds one sample and, of each 16-feature
// tile t, the features 16t + 4g + r.  The reduction index of the next
// product is permuted consistently in A and B (logical k = 8g + e of K-tile
// q <-> feature 16(2q + e/4) + 4g + e%4), so the D registers of one layer
// are the B operand of the next without any data movement.
#include <hip/hip_runtime.h>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>
#include "chanles.h"
#include "kernels.hpp"
#include "mlp.hpp"
#include "prof.hpp"
#include "wallmodel.hpp"

namespace chan {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float silu(float z) { return z / (1.0f + expf(-z)); }
__device__ __forceinline__ float dsilu(float z)
{
    const float s = 1.0f / (1.0f + expf(-z));
    return s * (1.0f + z * (1.0f - s));
}
// SiLU from the hardware's base-2 exponential and reciprocal (v_exp_f32,
// v_rcp_f32; about 1 ulp each, plus the rounding of -z log2(e), a relative
// 1e-6 for |z| < 20): five instructions where the library's correctly
// rounded expf and division take about twenty.  For the wall model in the
// step only: the activations are rounded to BF16 (8 bits) before the next
// layer, and the ~5x larger error here moves a few more of them across a
// rounding boundary than the library form does, so the network functions
// that are compared with the FP64 oracle sample by sample (inference,
// training) keep the library form.  z -> -inf gives exp2 = inf and 0.
__device__ __forceinline__ float silu_fast(float z)
{
    return z * __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(-1.4426950408889634f * z));
}
template <bool FAST>
__device__ __forceinline__ float silu_sel(float z)
{
    return FAST ? silu_fast(z) : silu(z);
}
__device__ __forceinline__ float bfr(float x) { return (float)(__bf16)x; }

// The network as the inference kernels read it: the forward MFMA A operands
// pre-permuted so that each lane reads its 8 values with one 16-byte load,
//   fw[L][m][q][lane]: hidden layer L (0: W2, 1: W3), A[n = 16m + l%16][k] = W[k][n],
// the input and output layers, the biases and the standardisation.  Built in
// global memory once per parameter change (k_mlp_image); inference blocks
// copy it into LDS with 16-byte loads.
struct alignas(16) MlpNet {
    bf16x8 fw[2][4][2][64];
    float w1[4 * 64], b1[64], b2[64], b3[64], w4[64 * 2], b4[2];
    float xm[4], xs[4], ym[2], ys[2];
};
static_assert(sizeof(MlpNet) % 16 == 0, "the image is copied in 16-byte pieces");
// training also holds the backward operands, A[k = 16m + l%16][n] = W[k][n]
struct MlpLds : MlpNet {
    bf16x8 bw[2][4][2][64];
};

__device__ __forceinline__ int perm_feature(int q, int g, int e) { return 16 * (2 * q + (e >> 2)) + 4 * g + (e & 3); }

// the network from its FP32 parameters p and standardisation norm, by the
// threads of one block (the backward operands when bw is given)
__device__ void mlp_fill(MlpNet& s, bf16x8 (*bw)[4][2][64], const float* __restrict__ p,
                         const float* __restrict__ norm)
{
    const int tid = threadIdx.x, nt = blockDim.x;
    for (int idx = tid; idx < 2 * 4 * 2 * 64; idx += nt) {
        const int lane = idx % 64, q = (idx / 64) % 2, m = (idx / 128) % 4, L = idx / 512;
        const float* W = p + (L == 0 ? MLP_W2 : MLP_W3);
        const int row = 16 * m + (lane % 16), g = lane / 16;
        bf16x8 f, b;
        for (int e = 0; e < 8; e++) {
            const int k = perm_feature(q, g, e);
            f[e] = (__bf16)W[k * 64 + row];   // W^T[row][k]
            b[e] = (__bf16)W[row * 64 + k];   // W[row][k]
        }
        s.fw[L][m][q][lane] = f;
        if (bw) bw[L][m][q][lane] = b;
    }
    for (int idx = tid; idx < 256; idx += nt) s.w1[idx] = bfr(p[MLP_W1 + idx]);
    for (int idx = tid; idx < 128; idx += nt) s.w4[idx] = bfr(p[MLP_W4 + idx]);
    for (int idx = tid; idx < 64; idx += nt) {
        s.b1[idx] = p[MLP_B1 + idx];
        s.b2[idx] = p[MLP_B2 + idx];
        s.b3[idx] = p[MLP_B3 + idx];
    }
    if (tid < 2) s.b4[tid] = p[MLP_B4 + tid];
    if (tid < 4) { s.xm[tid] = norm[tid]; s.xs[tid] = norm[4 + tid]; }
    if (tid < 2) { s.ym[tid] = norm[8 + tid]; s.ys[tid] = norm[10 + tid]; }
}

__device__ void mlp_load_lds(MlpLds& s, const float* __restrict__ p, const float* __restrict__ norm)
{
    mlp_fill(s, s.bw, p, norm);
}

// one block: the inference image of the parameters
__global__ void __launch_bounds__(256) k_mlp_image(const float* __restrict__ p, const float* __restrict__ norm,
                                                   MlpNet* __restrict__ img)
{
    mlp_fill(*img, nullptr, p, norm);
}

// the image into a block's LDS
__device__ void mlp_copy_image(MlpNet& s, const MlpNet* __restrict__ img)
{
    const uint4* __restrict__ src = reinterpret_cast<const uint4*>(img);
    uint4* dst = reinterpret_cast<uint4*>(&s);
    for (int q = threadIdx.x; q < (int)(sizeof(MlpNet) / 16); q += blockDim.x) dst[q] = src[q];
}

// Compiler-only fence at the top of a grid-stride loop over sample groups:
// the network image in LDS is loop-invariant, and without it the compiler
// keeps every weight a lane reads (~200 values) in registers across the
// loop, which holds the kernel to one wave per SIMD.  Re-reading them from
// LDS per group costs little and frees the registers for occupancy.
__device__ __forceinline__ void lds_reload_fence() { asm volatile("" ::: "memory"); }

// activations of one sample held by a lane: a[4t + r] <-> feature 16t + 4g + r
struct Act { float a[16]; };

__device__ __forceinline__ void mfma_layer(const bf16x8 (*A)[2][64], const Act& in, Act& out, int lane)
{
    bf16x8 bq[2];
    for (int q = 0; q < 2; q++)
        for (int e = 0; e < 8; e++) bq[q][e] = (__bf16)in.a[(2 * q + (e >> 2)) * 4 + (e & 3)];
    for (int m = 0; m < 4; m++) {
        floatx4 acc = {0.f, 0.f, 0.f, 0.f};
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[m][0][lane], bq[0], acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[m][1][lane], bq[1], acc, 0, 0, 0);
        for (int r = 0; r < 4; r++) out.a[m * 4 + r] = acc[r];
    }
}

// Forward pass of one 16-sample group.  x: standardised features of this
// lane's sample.  z1..z3: pre-activations; h3q: last hidden layer after BF16
// rounding; y: standardised outputs (every lane of a sample holds them).
template <bool FAST = false>
__device__ __forceinline__ void mlp_forward(const MlpNet& s, const float x[4], Act& z1, Act& z2, Act& z3, Act& h3q,
                                            float y[2])
{
    const int lane = threadIdx.x % 64, g = lane / 16;
    Act h;
    for (int t = 0; t < 4; t++)
        for (int r = 0; r < 4; r++) {
            const int n = 16 * t + 4 * g + r;
            float z = s.b1[n];
            for (int k = 0; k < 4; k++) z += x[k] * s.w1[k * 64 + n];
            z1.a[t * 4 + r] = z;
            h.a[t * 4 + r] = silu_sel<FAST>(z);
        }
    mfma_layer(s.fw[0], h, z2, lane);
    for (int q = 0; q < 16; q++) {
        const int n = 16 * (q / 4) + 4 * g + (q % 4);
        z2.a[q] += s.b2[n];
        h.a[q] = silu_sel<FAST>(z2.a[q]);
    }
    mfma_layer(s.fw[1], h, z3, lane);
    float y0 = 0.f, y1 = 0.f;
    for (int q = 0; q < 16; q++) {
        const int n = 16 * (q / 4) + 4 * g + (q % 4);
        z3.a[q] += s.b3[n];
        h3q.a[q] = bfr(silu_sel<FAST>(z3.a[q]));
        y0 += h3q.a[q] * s.w4[n * 2 + 0];
        y1 += h3q.a[q] * s.w4[n * 2 + 1];
    }
    // sum over the four lane groups of this sample, in a fixed order
    y0 += __shfl_xor(y0, 16);
    y1 += __shfl_xor(y1, 16);
    y0 += __shfl_xor(y0, 32);
    y1 += __shfl_xor(y1, 32);
    y[0] = y0 + s.b4[0];
    y[1] = y1 + s.b4[1];
}

// --------------------------------------------------------------- inference
// raw features in, raw outputs out (standardisation inside)
__global__ void __launch_bounds__(256) k_mlp_infer(const MlpNet* __restrict__ img, const float* __restrict__ X,
                                                   float* __restrict__ Y, int n)
{
    __shared__ MlpNet s;
    mlp_copy_image(s, img);
    __syncthreads();
    const int lane = threadIdx.x % 64, wave = threadIdx.x / 64;
    const int ngroups = (n + 15) / 16;
    for (int grp = blockIdx.x * 4 + wave; grp < ngroups; grp += gridDim.x * 4) {
        lds_reload_fence();
        const int smp = grp * 16 + lane % 16;
        const bool ok = smp < n;
        float x[4];
        for (int k = 0; k < 4; k++) x[k] = ok ? (X[(long)smp * 4 + k] - s.xm[k]) / s.xs[k] : 0.f;
        Act z1, z2, z3, h3q;
        float y[2];
        mlp_forward(s, x, z1, z2, z3, h3q, y);
        if (ok && lane < 16) {
            Y[(long)smp * 2 + 0] = y[0] * s.ys[0] + s.ym[0];
            Y[(long)smp * 2 + 1] = y[1] * s.ys[1] + s.ym[1];
        }
    }
}

// ---------------------------------------------------------------- training
// Forward + backward of one minibatch of B samples (B a multiple of 32;
// samples past nvalid contribute nothing).  Loss L = 1/(2 nvalid) sum over
// samples and both outputs of (y - t)^2 on standardised outputs.  Writes the
// per-sample quantities the weight gradients need, feature-major [f][B]:
//   Xt (4), H1 H2 H3 (BF16-rounded activations), DZ1 DZ2 DZ3 (pre-activation
//   gradients), DY (2); and one loss partial per 16-sample group.
struct TrainBufs {
    float *Xt, *H[3], *DZ[3], *DY, *loss;
};

__global__ void __launch_bounds__(256) k_mlp_fwd_bwd(const float* __restrict__ p, const float* __restrict__ norm,
                                                     const float* __restrict__ X, const float* __restrict__ Tg,
                                                     const int* __restrict__ idx, int nvalid, int B, TrainBufs tb)
{
    __shared__ MlpLds s;
    mlp_load_lds(s, p, norm);
    __syncthreads();
    const int lane = threadIdx.x % 64, wave = threadIdx.x / 64, g = lane / 16;
    const int ngroups = B / 16;
    const float inv = 1.0f / (float)nvalid;
    for (int grp = blockIdx.x * 4 + wave; grp < ngroups; grp += gridDim.x * 4) {
        lds_reload_fence();
        const int sl = grp * 16 + lane % 16;  // slot in the batch
        const bool ok = sl < nvalid;
        const long src = ok ? (long)idx[sl] : 0;
        float x[4], t[2];
        for (int k = 0; k < 4; k++) x[k] = ok ? (X[src * 4 + k] - s.xm[k]) / s.xs[k] : 0.f;
        for (int m = 0; m < 2; m++) t[m] = ok ? (Tg[src * 2 + m] - s.ym[m]) / s.ys[m] : 0.f;
        Act z1, z2, z3, h3q;
        float y[2];
        mlp_forward(s, x, z1, z2, z3, h3q, y);
        float dy[2], l2 = 0.f;
        for (int m = 0; m < 2; m++) {
            const float e = ok ? y[m] - t[m] : 0.f;
            dy[m] = e * inv;
            l2 += e * e;
        }
        // backward through the output layer and the SiLUs
        Act dz3, dz2, dz1, dh;
        for (int q = 0; q < 16; q++) {
            const int n = 16 * (q / 4) + 4 * g + (q % 4);
            dz3.a[q] = (s.w4[n * 2 + 0] * dy[0] + s.w4[n * 2 + 1] * dy[1]) * dsilu(z3.a[q]);
        }
        mfma_layer(s.bw[1], dz3, dh, lane);  // dH2 = W3 dZ3
        for (int q = 0; q < 16; q++) dz2.a[q] = dh.a[q] * dsilu(z2.a[q]);
        mfma_layer(s.bw[0], dz2, dh, lane);  // dH1 = W2 dZ2
        for (int q = 0; q < 16; q++) dz1.a[q] = dh.a[q] * dsilu(z1.a[q]);
        // per-sample quantities, feature-major
        for (int q = 0; q < 16; q++) {
            const int n = 16 * (q / 4) + 4 * g + (q % 4);
            const long o = (long)n * B + sl;
            tb.H[0][o] = bfr(silu(z1.a[q]));
            tb.H[1][o] = bfr(silu(z2.a[q]));
            tb.H[2][o] = h3q.a[q];
            tb.DZ[0][o] = ok ? dz1.a[q] : 0.f;
            tb.DZ[1][o] = ok ? dz2.a[q] : 0.f;
            tb.DZ[2][o] = ok ? dz3.a[q] : 0.f;
        }
        if (lane < 16) {
            for (int k = 0; k < 4; k++) tb.Xt[(long)k * B + sl] = x[k];
            tb.DY[sl] = dy[0];
            tb.DY[(long)B + sl] = dy[1];
        }
        // loss partial of this group: lanes 0..15 hold one sample each
        float lsum = lane < 16 ? l2 : 0.f;
        for (int o = 8; o >= 1; o >>= 1) lsum += __shfl_xor(lsum, o);
        if (lane == 0) tb.loss[grp] = 0.5f * lsum * inv;
    }
}

// Weight gradients.  Blocks 0..7: one wave per 16x16 tile of dW2 / dW3 =
// sum_s H_{L-1}[k][s] dZ_L[n][s] on the matrix cores (operands rounded to
// BF16, K = 32 samples per instruction).  Blocks 8..: one wave per small
// parameter (W1, b1..b3, W4, b4; 578 of them), its lanes striding over the
// samples (coalesced rows of the feature-major buffers), then a butterfly
// over the lanes.  Every sum runs in a fixed order.
constexpr int MLP_SMALL = 578;
constexpr int MLP_WGRAD_BLOCKS = 8 + (MLP_SMALL + 3) / 4;
__global__ void __launch_bounds__(256) k_mlp_wgrad(TrainBufs tb, int B, float* __restrict__ grad)
{
    const int lane = threadIdx.x % 64, wave = threadIdx.x / 64, g = lane / 16;
    if (blockIdx.x < 8) {
        const int tile = blockIdx.x * 4 + wave;  // 0..31
        const int L = tile / 16, mk = (tile / 4) % 4, mn = tile % 4;
        const float* H = tb.H[L];       // H1 for dW2, H2 for dW3
        const float* DZ = tb.DZ[L + 1]; // dZ2, dZ3
        const int k = 16 * mk + lane % 16, n = 16 * mn + lane % 16;
        floatx4 acc = {0.f, 0.f, 0.f, 0.f};
        for (int c = 0; c < B / 32; c++) {
            bf16x8 a, b;
            const long s0 = (long)c * 32 + 8 * g;
            for (int e = 0; e < 8; e++) {
                a[e] = (__bf16)H[(long)k * B + s0 + e];
                b[e] = (__bf16)DZ[(long)n * B + s0 + e];
            }
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc, 0, 0, 0);
        }
        float* dW = grad + (L == 0 ? MLP_W2 : MLP_W3);
        for (int r = 0; r < 4; r++) dW[(16 * mk + 4 * g + r) * 64 + n] = acc[r];
        return;
    }
    // small parameters: 256 (W1) + 64 x 3 (biases) + 128 (W4) + 2 (b4)
    const int q = (blockIdx.x - 8) * 4 + wave;
    if (q >= MLP_SMALL) return;
    const float *A = nullptr, *Bv = nullptr;  // the sum of A[s] * Bv[s] (Bv null: of A[s])
    float* out;
    if (q < 256) {
        const int k = q / 64, n = q % 64;
        A = tb.DZ[0] + (long)n * B;
        Bv = tb.Xt + (long)k * B;
        out = grad + MLP_W1 + q;
    } else if (q < 448) {
        const int L = (q - 256) / 64, n = (q - 256) % 64;
        A = tb.DZ[L] + (long)n * B;
        out = grad + (L == 0 ? MLP_B1 : (L == 1 ? MLP_B2 : MLP_B3)) + n;
    } else if (q < 576) {
        const int n = (q - 448) / 2, m = (q - 448) % 2;
        A = tb.H[2] + (long)n * B;
        Bv = tb.DY + (long)m * B;
        out = grad + MLP_W4 + q - 448;
    } else {
        A = tb.DY + (long)(q - 576) * B;
        out = grad + MLP_B4 + q - 576;
    }
    float sum = 0.f;
    if (Bv) {
        for (int s = lane; s < B; s += 64) sum += A[s] * Bv[s];
    } else {
        for (int s = lane; s < B; s += 64) sum += A[s];
    }
    for (int o = 32; o >= 1; o >>= 1) sum += __shfl_xor(sum, o);
    if (lane == 0) *out = sum;
}

// Adam (Kingma & Ba 2015) on the FP32 master parameters; c1 = 1/(1-b1^t),
// c2 = 1/(1-b2^t) from the host.
__global__ void k_mlp_adam(float* __restrict__ p, float* __restrict__ m, float* __restrict__ v,
                           const float* __restrict__ grad, float lr, float b1, float b2, float eps, float c1, float c2)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= MLP_NPARAM) return;
    const float gq = grad[q];
    const float mq = b1 * m[q] + (1.0f - b1) * gq;
    const float vq = b2 * v[q] + (1.0f - b2) * gq * gq;
    m[q] = mq;
    v[q] = vq;
    p[q] = p[q] - lr * (mq * c1) / (sqrtf(vq * c2) + eps);
}

// --------------------------------------------------------- the wall model
// Features of a wall cell (dimensionless, before standardisation):
//   f0 = asinh(h |u_par| / nu / 10)          local Reynolds number
//   f1 = asinh(h^2 d|u_par|/dn / nu / 10)    wall-normal gradient
//   f2 = asinh(y+_prev / 10), y+_prev = h sqrt(|tau_prev|) / nu
//   f3 = asinh(h^3 dpdx / nu^2 / 10)         driving pressure gradient
// Outputs: the stress along and across u_par,
//   o0 = log(1 + h^2 tau_par / nu^2),  o1 = h^2 tau_perp / nu^2 / exp(o0),
// rotated back to (tau_wx, tau_wz).  o0 = log(1 + y+^2) keeps the relative
// precision of tau_par nearly uniform from the sublayer to the log region,
// and o1 measures tau_perp relative to the same scale.  The equilibrium log law gives
// tau_par = u_tau(|u_par|)^2, tau_perp = 0: it is the training target, the
// yardstick of the network's error and the fallback.
__device__ __forceinline__ void wm_features(const WallSample& ws, double nu, double tau_prev, double dpdx, float f[4])
{
    const double up = sqrt(ws.u * ws.u + ws.w * ws.w);
    const double h = ws.h;
    f[0] = (float)asinh(h * up / nu / 10.0);
    f[1] = (float)asinh(h * h * ws.dudy / nu / 10.0);
    f[2] = (float)asinh(h * sqrt(fabs(tau_prev)) / nu / 10.0);
    f[3] = (float)asinh(h * h * h * dpdx / (nu * nu) / 10.0);
}

__device__ __forceinline__ double wm_driving_dpdx(const Scalars* sc, int x_mass_cte)
{
    // constant mass flow: the per-step correction U += last_dPdx acts as
    // last_dPdx / dt; otherwise the imposed dPdx
    if (x_mass_cte == 1 || x_mass_cte == 2) return sc->dt > 0.0 ? sc->last_dPdx / sc->dt : 0.0;
    return sc->dPdx;
}

// Training/validation samples of every wall cell at height h_wm, viscosity
// nu (features and the equilibrium targets o0, o1; the lagged stress is the
// context's alpha_x) and the equilibrium stress u_tau^2 itself (Tq).
template <typename T>
__global__ void k_wm_samples(const T* U, const T* W, const double* alpha, const Scalars* sc, Metric<double> m, Dims d,
                             double nu, double h_wm, double kappa, double B, double yc, int x_mass_cte, float* X,
                             float* Y, double* Tq)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int k = blockIdx.y + 1, wall = blockIdx.z;
    if (i > d.nx) return;
    const WallSample ws = wall_sample(U, W, m, d, i, k, wall, h_wm);
    const long c = (i - 1) + (long)d.nx * (wall + 2L * (k - 1));
    float f[4];
    wm_features(ws, nu, alpha[c], wm_driving_dpdx(sc, x_mass_cte), f);
    const double ut = loglaw_utau(sqrt(ws.u * ws.u + ws.w * ws.w), ws.h, nu, kappa, B, yc);
    for (int q = 0; q < 4; q++) X[c * 4 + q] = f[q];
    Y[c * 2 + 0] = (float)log1p(ws.h * ws.h * ut * ut / (nu * nu));
    Y[c * 2 + 1] = 0.0f;
    Tq[c] = ut * ut;
}

// The MLP wall model of one substep: per wall cell, the network's tau_wx
// unless its stress vector (tau_par, tau_perp) differs from the equilibrium
// log law's (u_tau^2, 0) by more than max_err relative to u_tau^2; then the
// equilibrium value.  Writes alpha_x and, per 16-cell group, the sum of
// relative errors and the number of fallbacks.
// A wave takes 64 cells at a time, one per lane: each lane samples the flow
// and forms the features of its own cell, the network runs on the four
// 16-cell groups in turn (an MFMA pass takes 16 samples, the features of
// group q broadcast from lanes 16q..16q+15), and each lane keeps the outputs
// of its own cell for the FP64 closure that follows.
template <typename T, bool G>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) k_wm_mlp(WmArgs<T> one, const WmArgs<T>* tab, Dims d)
{
    const GBlk gb = gblk(1);
    const WmArgs<T> arg = member_args<G>(one, tab, gb.m);
    const T *U = arg.U, *W = arg.W;
    double* alpha = arg.alpha;
    const Scalars* sc = arg.sc;
    const Metric<double>& m = arg.m;
    const double nu = arg.nu, h_wm = arg.h_wm, kappa = arg.kappa, B = arg.B, yc = arg.yc, max_err = arg.max_err;
    double* stats = arg.stats;
    const int x_mass_cte = arg.x_mass_cte;
    __shared__ MlpNet s;
    mlp_copy_image(s, static_cast<const MlpNet*>(arg.net));
    __syncthreads();
    const int lane = threadIdx.x % 64, wave = threadIdx.x / 64;
    const long ncell = 2L * d.nx * d.nzg;
    const long ngroups = (ncell + 15) / 16;
    const long nquad = (ncell + 63) / 64;
    const double dpdx = wm_driving_dpdx(sc, x_mass_cte);
    for (long qd = (long)gb.x * 4 + wave; qd < nquad; qd += (long)gb.gx * 4) {
        lds_reload_fence();
        const long c = qd * 64 + lane;
        const bool ok = c < ncell;
        float xo[4] = {0.f, 0.f, 0.f, 0.f};
        WallSample ws{0.0, 0.0, 0.0, 1.0};
        if (ok) {
            const int i = (int)(c % d.nx) + 1, wall = (int)((c / d.nx) % 2), k = (int)(c / (2L * d.nx)) + 1;
            ws = wall_sample(U, W, m, d, i, k, wall, h_wm);
            float f[4];
            wm_features(ws, nu, alpha[c], dpdx, f);
            for (int q = 0; q < 4; q++) xo[q] = (f[q] - s.xm[q]) / s.xs[q];
        }
        float ys[2] = {0.f, 0.f};
#pragma unroll 1
        for (int q = 0; q < 4; q++) {
            lds_reload_fence();
            float x[4];
            for (int e = 0; e < 4; e++) x[e] = __shfl(xo[e], 16 * q + lane % 16);
            Act z1, z2, z3, h3q;
            float y[2];
            mlp_forward<true>(s, x, z1, z2, z3, h3q, y);
            if (lane / 16 == q) { ys[0] = y[0]; ys[1] = y[1]; }
        }
        double err = 0.0, fb = 0.0;
        if (ok) {
            const double up = sqrt(ws.u * ws.u + ws.w * ws.w);
            const double cs = up > 0.0 ? ws.u / up : 1.0, sn = up > 0.0 ? ws.w / up : 0.0;
            const double sc2 = nu * nu / (ws.h * ws.h);
            const double o0 = (double)(ys[0] * s.ys[0] + s.ym[0]);
            const double tpar = sc2 * expm1(o0);
            const double tper = sc2 * exp(o0) * (double)(ys[1] * s.ys[1] + s.ym[1]);
            const double tx = tpar * cs - tper * sn;
            // the network's own u_tau starts the equilibrium Newton solve
            const double ut = loglaw_utau(up, ws.h, nu, kappa, B, yc, tpar > 0.0 ? sqrt(tpar) : 0.0);
            const double teq = ut * ut;
            const double tx_eq = teq * cs;
            // deviation of the whole wall-parallel stress vector
            err = teq > 0.0 ? sqrt((tpar - teq) * (tpar - teq) + tper * tper) / teq : 0.0;
            const bool use_eq = !(err <= max_err);
            fb = use_eq ? 1.0 : 0.0;
            alpha[c] = use_eq ? tx_eq : tx;
        }
        // sums over each 16-cell group, in a fixed order
        for (int o = 8; o >= 1; o >>= 1) {
            err += __shfl_xor(err, o);
            fb += __shfl_xor(fb, o);
        }
        const long grp = qd * 4 + lane / 16;
        if (lane % 16 == 0 && grp < ngroups) {
            stats[grp] = err;
            stats[ngroups + grp] = fb;
        }
    }
}

// ------------------------------------------------------------- launchers
void launch_mlp_infer(const void* img, const float* X, float* Y, int n, hipStream_t s)
{
    const int groups = (n + 15) / 16;
    const int blocks = std::max(1, std::min(1024, (groups + 3) / 4));
    PROF_BEGIN("k_mlp_infer", s);
    k_mlp_infer<<<blocks, 256, 0, s>>>(static_cast<const MlpNet*>(img), X, Y, n);
    PROF_END(s);
}

template <typename T>
void launch_wm_samples(const T* U, const T* W, const double* alpha, const Scalars* sc, const Metric<double>& m,
                       const Dims& d, double nu, double h_wm, double kappa, double B, int x_mass_cte, float* X, float* Y,
                       double* Tq, hipStream_t s)
{
    dim3 blk(64), grd(cdiv(d.nx, 64), d.nzg, 2);
    k_wm_samples<T><<<grd, blk, 0, s>>>(U, W, alpha, sc, m, d, nu, h_wm, kappa, B, loglaw_crossover(kappa, B),
                                         x_mass_cte, X, Y, Tq);
}

long wm_mlp_groups(const Dims& d) { return (2L * d.nx * d.nzg + 15) / 16; }

// blocks of 4 waves, one 64-cell set per wave: a set's forward passes take
// several microseconds of dependent instructions, so waves rather than
// loop trips carry the cells (the 18.8 KB image copy per block comes from L2)
template <typename T>
void run_wm_mlp(const WmArgs<T>& one, const WmArgs<T>* tab, int nmem, const Dims& d, hipStream_t s)
{
    const long quads = (2L * d.nx * d.nzg + 63) / 64;
    const int blocks = (int)std::max(1L, std::min(2048L, (quads + 3) / 4));
    PROF_BEGIN("k_wm_mlp", s);
    if (tab) k_wm_mlp<T, true><<<dim3(blocks, 1, nmem), 256, 0, s>>>(one, tab, d);
    else k_wm_mlp<T, false><<<dim3(blocks, 1, 1), 256, 0, s>>>(one, nullptr, d);
    PROF_END(s);
}

template <typename T>
void launch_wm_mlp(const void* net, const T* U, const T* W, double* alpha, const Scalars* sc,
                   const Metric<double>& m, const Dims& d, double nu, double h_wm, double kappa, double B,
                   int x_mass_cte, double max_err, double* stats, hipStream_t s)
{
    run_wm_mlp<T>(WmArgs<T>{net, U, W, alpha, sc, m, nu, h_wm, kappa, B, loglaw_crossover(kappa, B), max_err, stats,
                            x_mass_cte},
                  nullptr, 1, d, s);
}

template void launch_wm_samples<double>(const double*, const double*, const double*, const Scalars*,
                                        const Metric<double>&, const Dims&, double, double, double, double, int,
                                        float*, float*, double*, hipStream_t);
template void launch_wm_samples<float>(const float*, const float*, const double*, const Scalars*,
                                       const Metric<double>&, const Dims&, double, double, double, double, int, float*,
                                       float*, double*, hipStream_t);
template void run_wm_mlp<double>(const WmArgs<double>&, const WmArgs<double>*, int, const Dims&, hipStream_t);
template void run_wm_mlp<float>(const WmArgs<float>&, const WmArgs<float>*, int, const Dims&, hipStream_t);
template void launch_wm_mlp<double>(const void*, const double*, const double*, double*, const Scalars*,
                                    const Metric<double>&, const Dims&, double, double, double, double, int, double,
                                    double*, hipStream_t);
template void launch_wm_mlp<float>(const void*, const float*, const float*, double*, const Scalars*,
                                   const Metric<double>&, const Dims&, double, double, double, double, int, double,
                                   double*, hipStream_t);

}  // namespace chan

// ------------------------------------------------------------------ C ABI
using namespace chan;

// A network and the streams that read it.  Channels that use it for their
// wall model attach their stream.  Every parameter change is first ordered
// after the work already queued on those streams (an event recorded there,
// waited on by the network's stream), and ends with the rebuilt inference
// image and the `ready` event, which channels wait on before their next
// launch: a change never overlaps a step that reads the network.
struct chan_mlp {
    std::string err;
    int device = 0;
    hipStream_t stream = nullptr;
    float *p = nullptr, *m = nullptr, *v = nullptr, *grad = nullptr, *norm = nullptr;
    MlpNet* img = nullptr;
    hipEvent_t ready = nullptr;
    std::mutex mu;
    struct User {
        hipStream_t s;
        hipEvent_t ev;
    };
    std::vector<User> users;
    int adam_t = 0;
    // training buffers (grown on demand)
    int cap_B = 0;
    TrainBufs tb{};
    float* tbuf = nullptr;
    int* idx = nullptr;
    int cap_n = 0;
    float *X = nullptr, *Y = nullptr;
};

namespace chan {
bool mlp_device_net(const chan_mlp* net, const void** img, int* device)
{
    if (!net || !net->img) return false;
    *img = net->img;
    *device = net->device;
    return true;
}

hipEvent_t mlp_ready(const chan_mlp* net) { return net ? net->ready : nullptr; }

hipError_t mlp_attach(chan_mlp* net, hipStream_t s)
{
    std::lock_guard<std::mutex> lk(net->mu);
    for (const auto& u : net->users)
        if (u.s == s) return hipSuccess;
    hipEvent_t ev = nullptr;
    hipError_t e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
    if (e == hipSuccess) net->users.push_back({s, ev});
    return e;
}

void mlp_detach(chan_mlp* net, hipStream_t s)
{
    std::lock_guard<std::mutex> lk(net->mu);
    for (size_t q = 0; q < net->users.size(); q++)
        if (net->users[q].s == s) {
            (void)hipEventDestroy(net->users[q].ev);
            net->users.erase(net->users.begin() + q);
            return;
        }
}
}  // namespace chan

namespace {

int mfail(chan_mlp* c, int code, const std::string& msg)
{
    if (c) c->err = msg;
    return code;
}

#define MCHK(ctx, call)                                                                            \
    do {                                                                                           \
        hipError_t e_ = (call);                                                                    \
        if (e_ != hipSuccess)                                                                      \
            return mfail(ctx, CHAN_EHIP, std::string(#call) + ": " + hipGetErrorString(e_));      \
    } while (0)

uint64_t splitmix(uint64_t& s)
{
    uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
double unif(uint64_t& s) { return (double)(splitmix(s) >> 11) * (1.0 / 9007199254740992.0); }

int ensure_train(chan_mlp* c, int B, int n)
{
    if (B > c->cap_B) {
        if (c->tbuf) (void)hipFree(c->tbuf);
        if (c->idx) (void)hipFree(c->idx);
        c->tbuf = nullptr;
        c->idx = nullptr;
        const size_t per = (size_t)B * (4 + 6 * 64 + 2) + (size_t)(B / 16);
        MCHK(c, hipMalloc(&c->tbuf, per * sizeof(float)));
        MCHK(c, hipMalloc(&c->idx, (size_t)B * sizeof(int)));
        float* q = c->tbuf;
        c->tb.Xt = q; q += (size_t)4 * B;
        for (int L = 0; L < 3; L++) { c->tb.H[L] = q; q += (size_t)64 * B; }
        for (int L = 0; L < 3; L++) { c->tb.DZ[L] = q; q += (size_t)64 * B; }
        c->tb.DY = q; q += (size_t)2 * B;
        c->tb.loss = q;
        c->cap_B = B;
    }
    if (n > c->cap_n) {
        if (c->X) (void)hipFree(c->X);
        if (c->Y) (void)hipFree(c->Y);
        c->X = c->Y = nullptr;
        MCHK(c, hipMalloc(&c->X, (size_t)n * 4 * sizeof(float)));
        MCHK(c, hipMalloc(&c->Y, (size_t)n * 2 * sizeof(float)));
        c->cap_n = n;
    }
    return CHAN_OK;
}

int mset_device(chan_mlp* c)
{
    MCHK(c, hipSetDevice(c->device));
    return CHAN_OK;
}

// order the network's stream after the work queued on its users' streams
int begin_update(chan_mlp* c)
{
    std::lock_guard<std::mutex> lk(c->mu);
    for (const auto& u : c->users) {
        MCHK(c, hipEventRecord(u.ev, u.s));
        MCHK(c, hipStreamWaitEvent(c->stream, u.ev, 0));
    }
    return CHAN_OK;
}

// the inference image of the new parameters, then the ready event
int end_update(chan_mlp* c)
{
    k_mlp_image<<<1, 256, 0, c->stream>>>(c->p, c->norm, c->img);
    MCHK(c, hipGetLastError());
    MCHK(c, hipEventRecord(c->ready, c->stream));
    return CHAN_OK;
}

// one optimiser step on the batch idx[0..nvalid) (already on the device)
int train_batch(chan_mlp* c, int nvalid, int B, float lr)
{
    const int groups = B / 16;
    const int blocks = std::max(1, (groups + 3) / 4);
    k_mlp_fwd_bwd<<<blocks, 256, 0, c->stream>>>(c->p, c->norm, c->X, c->Y, c->idx, nvalid, B, c->tb);
    k_mlp_wgrad<<<MLP_WGRAD_BLOCKS, 256, 0, c->stream>>>(c->tb, B, c->grad);
    c->adam_t++;
    const double b1 = 0.9, b2 = 0.999;
    const float c1 = (float)(1.0 / (1.0 - std::pow(b1, c->adam_t)));
    const float c2 = (float)(1.0 / (1.0 - std::pow(b2, c->adam_t)));
    k_mlp_adam<<<cdiv(MLP_NPARAM, 256), 256, 0, c->stream>>>(c->p, c->m, c->v, c->grad, lr, (float)b1, (float)b2,
                                                             1e-8f, c1, c2);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return mfail(c, CHAN_EHIP, std::string("mlp training launch: ") + hipGetErrorString(e));
    return CHAN_OK;
}

}  // namespace

extern "C" {

int chan_mlp_create(int device, unsigned long long seed, chan_mlp** out)
{
    if (!out) return CHAN_EINVAL;
    chan_mlp* c = new chan_mlp();
    *out = c;
    c->device = device;
    int rc = mset_device(c);
    if (rc) return rc;
    MCHK(c, hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    MCHK(c, hipMalloc(&c->p, MLP_NPARAM * sizeof(float)));
    MCHK(c, hipMalloc(&c->m, MLP_NPARAM * sizeof(float)));
    MCHK(c, hipMalloc(&c->v, MLP_NPARAM * sizeof(float)));
    MCHK(c, hipMalloc(&c->grad, MLP_NPARAM * sizeof(float)));
    MCHK(c, hipMalloc(&c->norm, MLP_NNORM * sizeof(float)));
    MCHK(c, hipMalloc(&c->img, sizeof(MlpNet)));
    MCHK(c, hipEventCreateWithFlags(&c->ready, hipEventDisableTiming));
    // He-uniform weights (bound sqrt(6 / fan_in)), zero biases, identity
    // standardisation; seeded counter RNG
    std::vector<float> p(MLP_NPARAM, 0.0f);
    uint64_t st = seed;
    auto fill = [&](int off, int fan_in, int count) {
        const double bnd = std::sqrt(6.0 / fan_in);
        for (int q = 0; q < count; q++) p[off + q] = (float)((2.0 * unif(st) - 1.0) * bnd);
    };
    fill(MLP_W1, 4, 256);
    fill(MLP_W2, 64, 4096);
    fill(MLP_W3, 64, 4096);
    fill(MLP_W4, 64, 128);
    float norm[MLP_NNORM] = {0, 0, 0, 0, 1, 1, 1, 1, 0, 0, 1, 1};
    MCHK(c, hipMemcpy(c->p, p.data(), MLP_NPARAM * sizeof(float), hipMemcpyHostToDevice));
    MCHK(c, hipMemset(c->m, 0, MLP_NPARAM * sizeof(float)));
    MCHK(c, hipMemset(c->v, 0, MLP_NPARAM * sizeof(float)));
    MCHK(c, hipMemset(c->grad, 0, MLP_NPARAM * sizeof(float)));
    MCHK(c, hipMemcpy(c->norm, norm, sizeof(norm), hipMemcpyHostToDevice));
    int rc2 = end_update(c);
    if (rc2) return rc2;
    MCHK(c, hipStreamSynchronize(c->stream));
    return CHAN_OK;
}

int chan_mlp_destroy(chan_mlp* c)
{
    if (!c) return CHAN_OK;
    {
        std::lock_guard<std::mutex> lk(c->mu);
        if (!c->users.empty())
            return mfail(c, CHAN_ESTATE, "the network is the wall model of " + std::to_string(c->users.size()) +
                                             " channel(s); detach it first (chan_set_wall_network(c, NULL, ...))");
    }
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    float* bufs[] = {c->p, c->m, c->v, c->grad, c->norm, c->tbuf, c->X, c->Y};
    for (float* b : bufs)
        if (b) (void)hipFree(b);
    if (c->img) (void)hipFree(c->img);
    if (c->idx) (void)hipFree(c->idx);
    if (c->ready) (void)hipEventDestroy(c->ready);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return CHAN_OK;
}

const char* chan_mlp_last_error(const chan_mlp* c) { return c ? c->err.c_str() : "null network"; }

int chan_mlp_get_params(chan_mlp* c, float* params, float* norm)
{
    if (!c) return CHAN_EINVAL;
    int rc = mset_device(c);
    if (rc) return rc;
    MCHK(c, hipStreamSynchronize(c->stream));
    if (params) MCHK(c, hipMemcpy(params, c->p, MLP_NPARAM * sizeof(float), hipMemcpyDeviceToHost));
    if (norm) MCHK(c, hipMemcpy(norm, c->norm, MLP_NNORM * sizeof(float), hipMemcpyDeviceToHost));
    return CHAN_OK;
}

int chan_mlp_set_params(chan_mlp* c, const float* params, const float* norm)
{
    if (!c) return CHAN_EINVAL;
    if (norm) {
        for (int q = 4; q < 8; q++)
            if (!(norm[q] > 0.0f)) return mfail(c, CHAN_EINVAL, "feature scales must be positive");
        for (int q = 10; q < 12; q++)
            if (!(norm[q] > 0.0f)) return mfail(c, CHAN_EINVAL, "output scales must be positive");
    }
    int rc = mset_device(c);
    if (rc) return rc;
    if ((rc = begin_update(c))) return rc;
    if (norm) MCHK(c, hipMemcpyAsync(c->norm, norm, MLP_NNORM * sizeof(float), hipMemcpyHostToDevice, c->stream));
    if (params) {
        MCHK(c, hipMemcpyAsync(c->p, params, MLP_NPARAM * sizeof(float), hipMemcpyHostToDevice, c->stream));
        MCHK(c, hipMemsetAsync(c->m, 0, MLP_NPARAM * sizeof(float), c->stream));
        MCHK(c, hipMemsetAsync(c->v, 0, MLP_NPARAM * sizeof(float), c->stream));
        c->adam_t = 0;
    }
    if ((rc = end_update(c))) return rc;
    // the host arrays may be released on return
    MCHK(c, hipStreamSynchronize(c->stream));
    return CHAN_OK;
}

int chan_mlp_forward(chan_mlp* c, const float* X, int n, float* Y)
{
    if (!c || !X || !Y || n < 0) return CHAN_EINVAL;
    if (n == 0) return CHAN_OK;
    int rc = mset_device(c);
    if (rc) return rc;
    if ((rc = ensure_train(c, 32, n))) return rc;
    MCHK(c, hipMemcpyAsync(c->X, X, (size_t)n * 4 * sizeof(float), hipMemcpyHostToDevice, c->stream));
    launch_mlp_infer(c->img, c->X, c->Y, n, c->stream);
    MCHK(c, hipGetLastError());
    MCHK(c, hipMemcpyAsync(Y, c->Y, (size_t)n * 2 * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    MCHK(c, hipStreamSynchronize(c->stream));
    return CHAN_OK;
}

int chan_mlp_gradients(chan_mlp* c, const float* X, const float* Y, int n, float* grad, double* loss)
{
    if (!c || !X || !Y || !grad || n < 1) return CHAN_EINVAL;
    int rc = mset_device(c);
    if (rc) return rc;
    const int B = (n + 31) / 32 * 32;
    if ((rc = ensure_train(c, B, n))) return rc;
    std::vector<int> id(B);
    for (int q = 0; q < B; q++) id[q] = q < n ? q : 0;
    MCHK(c, hipMemcpyAsync(c->X, X, (size_t)n * 4 * sizeof(float), hipMemcpyHostToDevice, c->stream));
    MCHK(c, hipMemcpyAsync(c->Y, Y, (size_t)n * 2 * sizeof(float), hipMemcpyHostToDevice, c->stream));
    MCHK(c, hipMemcpyAsync(c->idx, id.data(), (size_t)B * sizeof(int), hipMemcpyHostToDevice, c->stream));
    const int groups = B / 16;
    k_mlp_fwd_bwd<<<std::max(1, (groups + 3) / 4), 256, 0, c->stream>>>(c->p, c->norm, c->X, c->Y, c->idx, n, B,
                                                                        c->tb);
    k_mlp_wgrad<<<MLP_WGRAD_BLOCKS, 256, 0, c->stream>>>(c->tb, B, c->grad);
    MCHK(c, hipGetLastError());
    std::vector<float> lp(groups);
    MCHK(c, hipMemcpyAsync(grad, c->grad, MLP_NPARAM * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    MCHK(c, hipMemcpyAsync(lp.data(), c->tb.loss, groups * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    MCHK(c, hipStreamSynchronize(c->stream));
    if (loss) {
        double s = 0.0;
        for (float v : lp) s += v;
        *loss = s;
    }
    return CHAN_OK;
}

int chan_mlp_fit_normalization(chan_mlp* c, const float* X, const float* Y, int n)
{
    if (!c || !X || !Y || n < 2) return CHAN_EINVAL;
    float norm[MLP_NNORM];
    for (int f = 0; f < 6; f++) {
        const float* src = f < 4 ? X : Y;
        const int stride = f < 4 ? 4 : 2, col = f < 4 ? f : f - 4;
        double s = 0.0, s2 = 0.0;
        for (int q = 0; q < n; q++) s += src[(long)q * stride + col];
        const double mean = s / n;
        for (int q = 0; q < n; q++) {
            const double dlt = src[(long)q * stride + col] - mean;
            s2 += dlt * dlt;
        }
        double sd = std::sqrt(s2 / n);
        if (!(sd > 1e-6)) sd = 1.0;
        if (f < 4) { norm[f] = (float)mean; norm[4 + f] = (float)sd; }
        else { norm[8 + col] = (float)mean; norm[10 + col] = (float)sd; }
    }
    return chan_mlp_set_params(c, nullptr, norm);
}

int chan_mlp_train(chan_mlp* c, const float* X, const float* Y, int n, int epochs, int batch, float lr,
                   unsigned long long seed, double* epoch_loss)
{
    if (!c || !X || !Y || n < 1 || epochs < 0 || batch < 32 || batch % 32 != 0 || !(lr > 0.0f)) {
        if (c) c->err = "mlp train: invalid arguments (batch must be a positive multiple of 32)";
        return CHAN_EINVAL;
    }
    int rc = mset_device(c);
    if (rc) return rc;
    if ((rc = ensure_train(c, batch, n))) return rc;
    if ((rc = begin_update(c))) return rc;
    MCHK(c, hipMemcpyAsync(c->X, X, (size_t)n * 4 * sizeof(float), hipMemcpyHostToDevice, c->stream));
    MCHK(c, hipMemcpyAsync(c->Y, Y, (size_t)n * 2 * sizeof(float), hipMemcpyHostToDevice, c->stream));
    std::vector<int> perm(n);
    for (int q = 0; q < n; q++) perm[q] = q;
    int* dperm = nullptr;
    MCHK(c, hipMalloc(&dperm, (size_t)n * sizeof(int)));
    const int nb = (n + batch - 1) / batch;
    float* dloss = nullptr;
    hipError_t e = hipMalloc(&dloss, (size_t)nb * (batch / 16) * sizeof(float));
    if (e != hipSuccess) {
        (void)hipFree(dperm);
        return mfail(c, CHAN_ENOMEM, "mlp train: loss buffer");
    }
    uint64_t st = seed;
    std::vector<float> lp((size_t)nb * (batch / 16));
    float* keep_loss = c->tb.loss;
    for (int ep = 0; ep < epochs && rc == CHAN_OK; ep++) {
        // Fisher-Yates shuffle with the seeded generator
        for (int q = n - 1; q > 0; q--) {
            const int r = (int)(splitmix(st) % (uint64_t)(q + 1));
            std::swap(perm[q], perm[r]);
        }
        e = hipMemcpyAsync(dperm, perm.data(), (size_t)n * sizeof(int), hipMemcpyHostToDevice, c->stream);
        if (e != hipSuccess) { rc = mfail(c, CHAN_EHIP, "mlp train: permutation upload"); break; }
        for (int b = 0; b < nb && rc == CHAN_OK; b++) {
            const int nvalid = std::min(batch, n - b * batch);
            // the batch's indices are a slice of the permutation: point idx at it
            int* keep_idx = c->idx;
            c->idx = dperm + (size_t)b * batch;
            c->tb.loss = dloss + (size_t)b * (batch / 16);
            rc = train_batch(c, nvalid, nvalid == batch ? batch : (nvalid + 31) / 32 * 32, lr);
            c->idx = keep_idx;
        }
        if (rc) break;
        e = hipMemcpyAsync(lp.data(), dloss, lp.size() * sizeof(float), hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) { rc = mfail(c, CHAN_EHIP, std::string("mlp train: ") + hipGetErrorString(e)); break; }
        if (epoch_loss) {
            // mean per-sample loss over the epoch: batch losses weighted by size
            double s = 0.0;
            for (int b = 0; b < nb; b++) {
                const int nvalid = std::min(batch, n - b * batch);
                double lb = 0.0;
                const int g = ((nvalid == batch ? batch : (nvalid + 31) / 32 * 32)) / 16;
                for (int q = 0; q < g; q++) lb += lp[(size_t)b * (batch / 16) + q];
                s += lb * nvalid;
            }
            epoch_loss[ep] = s / n;
        }
    }
    c->tb.loss = keep_loss;
    // the image of whatever the parameters now are, also after a failure
    const int rc_img = end_update(c);
    (void)hipStreamSynchronize(c->stream);
    (void)hipFree(dperm);
    (void)hipFree(dloss);
    return rc ? rc : rc_img;
}

}  // extern "C"
