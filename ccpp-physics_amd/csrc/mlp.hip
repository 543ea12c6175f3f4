// Learned LW gas optics: a 7 -> 128 -> 128 -> 128 -> 256 softplus MLP whose
// GEMMs run on gfx950 MFMA (v_mfma_f32_32x32x16_bf16: BF16 inputs, FP32
// accumulation).  It stands in for the k-distribution lookup of
// ty_gas_optics_rrtmgp%gas_optics (rrtmgp_lw_main.F90:337-345): per layer the
// network maps (T, ln p, ln vmr of H2O O3 CO2 CH4 N2O), standardised, to
// 128 values of ln(tau_g / col_dry) and 128 Planck fractions, standardised.
//
// MFMA 32x32x16 lane layout used throughout (checked by the layout test):
//   A (32 x 16): lane l holds A[l % 32][8*(l/32) + j], j = 0..7
//   B (16 x 32): lane l holds B[8*(l/32) + j][l % 32]
//   D (32 x 32): lane l, register r holds D[8*(r/4) + 4*(l/32) + r%4][l % 32]
//
// Training: mlp_gemm_kernel computes C = op(A) op(B) for arbitrary strides
// (forward X W, backward dY W^T and X^T dY) through 64 x 64 LDS tiles; the
// epilogue adds the bias and applies softplus (forward) or multiplies by
// sigmoid(Z) (backward).  Weight gradients use deterministic split-K.
//
// Inference: mlp_infer_kernel keeps the four weight matrices in LDS as ready
// A-fragments (transposed: W^T is the A operand, activations the B operand),
// so the D registers of one layer are, after bias + softplus + BF16 rounding,
// the B registers of the next: the next layer's K order is permuted to
// match, by storing its weights accordingly.  One wavefront carries one
// 32-row tile at a time; the hidden layers work in base 2 (softplus / ln2,
// one exp2 and one log2 per activation), the output layer's products for
// one pair of feature tiles overlap the previous pair's transform and
// stores, and the next tile's inputs are loaded during the current one.
#include "kernels.h"

#include <type_traits>

namespace ccpprad {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int MLP_IN = MLP_NIN, MLP_H = 128, MLP_OUT = 256, MLP_G = 128;
constexpr int MLP_K1 = 14;
constexpr double LOG2E_D = 1.4426950408889634, LN2_D = 0.6931471805599453;   // layer-1 K: hi and lo BF16 parts of the 7 inputs
constexpr int MLP_NPARAM = 7 * 128 + 128 + 2 * (128 * 128 + 128) + 128 * 256 + 256;  // 67072

__device__ __forceinline__ float softplus_f(float z)
{
    // max(z,0) + log1p(exp(-|z|)) with the hardware exp2/log2
    const float a = fabsf(z);
    const float e = __builtin_amdgcn_exp2f(-a * 1.4426950408889634f);
    return fmaxf(z, 0.0f) + __builtin_amdgcn_logf(1.0f + e) * 0.6931471805599453f;
}

// log2(1 + 2^z) = max(z, 0) + log2(1 + 2^-|z|), the hardware exp2/log2: the
// softplus of z ln2, divided by ln2
__device__ __forceinline__ float softplus2_f(float z)
{
    const float e = __builtin_amdgcn_exp2f(-fabsf(z));
    return fmaxf(z, 0.0f) + __builtin_amdgcn_logf(1.0f + e);
}

__device__ __forceinline__ float sigmoid_f(float z) { return 1.0f / (1.0f + __expf(-z)); }

__device__ __forceinline__ __bf16 f2bf(float f) { return (__bf16)f; }

// ---------------------------------------------------------------------------
// Training GEMM
// ---------------------------------------------------------------------------
struct GemmArgs {
    int M, N, K;                       // C is M x N, reduction over K
    const __bf16 *A; long sam, sak;    // A(m,k) = A[m*sam + k*sak]
    const __bf16 *B; long sbk, sbn;    // B(k,n) = B[k*sbk + n*sbn]
    const float *bias;                 // [N] or null
    const float *Z;                    // backward: multiply by sigmoid(Z[m*ldz + n])
    long ldz;
    float *C32; long ldc;              // fp32 output (or split-K partial base)
    __bf16 *C16; long ldc16;           // bf16 output (softplus applied if act)
    int act;                           // 1: softplus on the bf16 output
    int ksplit;                        // K range per blockIdx.z (split-K), 0 = all
};

__global__ __launch_bounds__(256) void mlp_gemm_kernel(GemmArgs g)
{
    __shared__ __attribute__((aligned(16))) __bf16 As[64][32];
    __shared__ __attribute__((aligned(16))) __bf16 Bs[64][32];   // n-major
    const int t = threadIdx.x, l = t & 63, w = t >> 6;
    const int wm = w >> 1, wn = w & 1;
    const int m0 = blockIdx.y * 64, n0 = blockIdx.x * 64;
    int kb = 0, ke = g.K;
    if (g.ksplit > 0) { kb = blockIdx.z * g.ksplit; ke = min(g.K, kb + g.ksplit); }
    f32x16 acc = {};
    const int lr = t >> 2, lk = (t & 3) * 8;
    for (int k0 = kb; k0 < ke; k0 += 32) {
        // cooperative tile loads with zero fill
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const int m = m0 + lr, k = k0 + lk + j;
            As[lr][lk + j] = (m < g.M && k < ke) ? g.A[(long)m * g.sam + (long)k * g.sak] : f2bf(0.0f);
            const int n = n0 + lr;
            Bs[lr][lk + j] = (n < g.N && k < ke) ? g.B[(long)k * g.sbk + (long)n * g.sbn] : f2bf(0.0f);
        }
        __syncthreads();
#pragma unroll
        for (int ks = 0; ks < 2; ++ks) {
            const bf16x8 a = *(const bf16x8 *)&As[wm * 32 + (l & 31)][ks * 16 + 8 * (l >> 5)];
            const bf16x8 b = *(const bf16x8 *)&Bs[wn * 32 + (l & 31)][ks * 16 + 8 * (l >> 5)];
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
        }
        __syncthreads();
    }
    const int n = n0 + wn * 32 + (l & 31);
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const int m = m0 + wm * 32 + 8 * (r >> 2) + 4 * (l >> 5) + (r & 3);
        if (m >= g.M || n >= g.N) continue;
        float v = acc[r];
        if (g.bias) v += g.bias[n];
        if (g.Z) v *= sigmoid_f(g.Z[(long)m * g.ldz + n]);
        if (g.C32) {
            float *cp = g.C32 + (g.ksplit > 0 ? (long)blockIdx.z * g.M * g.ldc : 0);
            cp[(long)m * g.ldc + n] = v;
        }
        if (g.C16) g.C16[(long)m * g.ldc16 + n] = f2bf(g.act ? softplus_f(v) : v);
    }
}

static void gemm(const GemmArgs &a, int splits, hipStream_t st)
{
    dim3 grid((a.N + 63) / 64, (a.M + 63) / 64, splits > 0 ? splits : 1);
    hipLaunchKernelGGL(mlp_gemm_kernel, grid, dim3(256), 0, st, a);
}

// sum of split-K partials: out[i] = sum_s part[s*n + i]
__global__ void sum_partials_kernel(const float *part, int splits, long n, float *out)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float s = 0.0f;
    for (int k = 0; k < splits; ++k) s += part[(long)k * n + i];
    out[i] = s;
}

// column sums (bias gradients): out[n] = sum_m x[m*ld + n]
__global__ void colsum_kernel(const float *x, int M, int N, long ld, float *out)
{
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    float s = 0.0f;
    for (int m = 0; m < M; ++m) s += x[(long)m * ld + n];
    out[n] = s;
}

// dY = 2 (Y - T) / (rows * 256): fp32 and bf16 copies; per-block loss partial
__global__ void mse_grad_kernel(const float *Y, const float *T, long n, float inv_n, float *dY32, __bf16 *dY16,
                                float *loss_part)
{
    __shared__ float red[256];
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    float d = 0.0f;
    if (i < n) {
        d = Y[i] - T[i];
        const float gval = 2.0f * d * inv_n;
        dY32[i] = gval;
        dY16[i] = f2bf(gval);
    }
    red[threadIdx.x] = d * d;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) loss_part[blockIdx.x] = red[0];
}

__global__ void sum_kernel(const float *x, int n, float scale, float *out)
{
    __shared__ float red[256];
    float s = 0.0f;
    for (int i = threadIdx.x; i < n; i += 256) s += x[i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if (threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = red[0] * scale;
}

// Adam on FP32 master parameters; refreshes the BF16 copies.
__global__ void adam_kernel(float *p, const float *g, float *m, float *v, __bf16 *p16, int n, float lr, float b1,
                            float b2, float eps, float c1, float c2)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float pi = p[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    p[i] = pi;
    p16[i] = f2bf(pi);
}

// standardised FP32 features (rows x 7) -> layer-1 input (rows x 14): BF16
// hi parts, then BF16 lo parts (x - hi), so the inputs keep ~16 bits.
__global__ void prep_x14_kernel(const float *x, int rows, __bf16 *x14)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long)rows * MLP_NIN) return;
    const long r = i / MLP_NIN;
    const int k = (int)(i % MLP_NIN);
    const float v = x[i];
    const __bf16 hi = (__bf16)v;
    x14[r * 14 + k] = hi;
    x14[r * 14 + MLP_NIN + k] = (__bf16)(v - (float)hi);
}

// [W1; W1] (14 x 128) from the BF16 W1 (7 x 128)
__global__ void w1eff_kernel(const __bf16 *w1, __bf16 *w1e)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 14 * 128) return;
    w1e[i] = w1[i % (7 * 128)];
}

// dW1 = dW1eff[0:7] + dW1eff[7:14]
__global__ void fold_w1_kernel(const float *g14, float *g7)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 7 * 128) return;
    g7[i] = g14[i] + g14[i + 7 * 128];
}

__global__ void to_bf16_kernel(const float *x, __bf16 *y, long n)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = f2bf(x[i]);
}

// ---------------------------------------------------------------------------
// Inference
// ---------------------------------------------------------------------------
// LDS image (bytes): A-fragments of W1^T (4 tiles x 1 chunk), W2^T (4 x 8),
// W3^T (4 x 8), W4^T (8 x 8): 1 KB each; then biases (FP32).
constexpr int FRAG_BYTES = 64 * 16;
constexpr int NFRAG = 4 * 1 + 4 * 8 + 4 * 8 + 8 * 8;              // 132
// fragments, then biases (3 x 128 + 256) and the output de-standardisation
// (ystd, ymean: 2 x 256), all FP32
constexpr int MLP_LDS_BYTES = NFRAG * FRAG_BYTES + (3 * 128 + 256 + 2 * 256) * 4;


// The raw inputs of one tile row (the 9 FP64 loads a lane makes for its row),
// fetched one tile ahead so their latency hides behind the current tile.
struct MlpRow {
    double t, p, v[5], pl0, pl1;
    long orow;      // output row, [col][lay] order
    int dcol;
    bool valid;
};

__device__ __forceinline__ void mlp_load_row(const MlpInferArgs &A, long nrows, long tile, int rl, MlpRow &R)
{
    const long row = tile * 32 + rl;
    R.valid = row < nrows;
    const long rr = R.valid ? row : nrows - 1;
    // rows are ordered column-fastest within a layer, so the 32 rows of a tile
    // read consecutive columns of the (nCol, nLay) inputs
    const int d = (int)(rr % A.ncols), lay = (int)(rr / A.ncols);
    const long q = (long)(A.col0 + d) + (long)A.ncol_ld * lay;
    R.orow = (long)d * A.nlay + lay;
    R.dcol = d;
    R.t = A.t_lay[q];
    R.p = A.p_lay[q];
#pragma unroll
    for (int i = 0; i < 5; ++i) R.v[i] = A.vmr[i][q];
    R.pl0 = A.p_lev[q];
    R.pl1 = A.p_lev[q + A.ncol_ld];
}

// The hidden layers run in the base-2 domain: with z' = z log2(e) the
// softplus is ln2 * softplus2(z'), so the standardised features and the
// biases of layers 1-3 carry a factor log2(e), each hidden activation is
// softplus2(z') (the next layer's weights then see activations 1/ln2 times
// the natural ones, which the factor log2(e) of its bias and pre-activation
// absorbs exactly), and the output scale ystd carries the factor ln2.
// One wavefront carries one 32-row tile at a time; 8 wavefronts per block
// share the LDS weight image.
__global__ __launch_bounds__(512) void mlp_infer_kernel(MlpInferArgs A)
{
    extern __shared__ __attribute__((aligned(16))) unsigned char lds[];
    const int t = threadIdx.x, l = t & 63, w = t >> 6;
    // stage the image (weights + biases) into LDS
    {
        const int4 *src = (const int4 *)A.image;
        int4 *dst = (int4 *)lds;
        for (int i = t; i < MLP_LDS_BYTES / 16; i += blockDim.x) dst[i] = src[i];
    }
    __syncthreads();
    const bf16x8 *F1 = (const bf16x8 *)lds;
    const bf16x8 *F2 = F1 + 4 * 64;
    const bf16x8 *F3 = F2 + 32 * 64;
    const bf16x8 *F4 = F3 + 32 * 64;
    const float *bias = (const float *)(lds + NFRAG * FRAG_BYTES);
    const float *ysd = bias + 640, *yc = bias + 896;
    const int half = l >> 5, rl = l & 31;
    const long nrows = (long)A.ncols * A.nlay;
    const long ntiles = (nrows + 31) / 32;
    const long stride = (long)gridDim.x * 8;
    long tile = (long)blockIdx.x * 8 + w;
    MlpRow cur;
    if (tile < ntiles) mlp_load_row(A, nrows, tile, rl, cur);
    for (; tile < ntiles; tile += stride) {
        // features (standardised, split into BF16 hi + lo) and dry-air column
        float f[MLP_NIN];
        f[0] = (float)cur.t;
        f[1] = __logf((float)cur.p);
#pragma unroll
        for (int i = 0; i < 5; ++i) f[2 + i] = __logf((float)cur.v[i]);
        bool oob = false;
#pragma unroll
        for (int i = 0; i < MLP_NIN; ++i) {
            oob |= !(f[i] >= A.lo[i] && f[i] <= A.hi[i]);
            f[i] = (f[i] - A.xmean[i]) * A.xistd[i];
        }
        const double vh = cur.v[0], fact = 1.0 / (1.0 + vh);
        const float col_dry = (float)(10.0 * fabs(cur.pl0 - cur.pl1) * AVOGAD /
                                      (1000.0 * ((M_DRY + M_H2O * vh) * fact) * 100.0 * GRAV) * fact);
        const long orow = cur.orow;
        const int dcol = cur.dcol;
        const bool valid = cur.valid;
        // the next tile's inputs, in flight during this tile's products
        if (tile + stride < ntiles) mlp_load_row(A, nrows, tile + stride, rl, cur);

        // layer 1: B operand slot k = hi_k (k < 7), lo_{k-7} (7 <= k < 14), 0
        bf16x8 h[8];   // B operands of the next layer (8 K-chunks of 16)
        {
            bf16x8 b1;
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const int k = 8 * half + j;
                float v = 0.0f;
#pragma unroll
                for (int i = 0; i < MLP_NIN; ++i) {
                    const float hi = (float)f2bf(f[i]);
                    if (k == i) v = hi;
                    if (k == i + MLP_NIN) v = (float)f2bf(f[i] - hi);
                }
                b1[j] = f2bf(v);
            }
#pragma unroll
            for (int np = 0; np < 2; ++np) {
                f32x16 a0 = {}, a1 = {};
                a0 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(F1[(2 * np) * 64 + l], b1, a0, 0, 0, 0);
                a1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(F1[(2 * np + 1) * 64 + l], b1, a1, 0, 0, 0);
#pragma unroll
                for (int r = 0; r < 16; ++r) {
                    const int n0 = (2 * np) * 32 + 8 * (r >> 2) + 4 * half + (r & 3);
                    h[4 * np + (r >> 3)][r & 7] = f2bf(softplus2_f(a0[r] + bias[n0]));
                    h[4 * np + 2 + (r >> 3)][r & 7] = f2bf(softplus2_f(a1[r] + bias[n0 + 32]));
                }
            }
        }
        // layers 2, 3: both halves of the output features' MFMA chains, then
        // their softplus, scheduled so that each of the second half's 16
        // products issues beside 12 VALU operations of the first half's
        // softplus
#pragma unroll 1
        for (int L = 0; L < 2; ++L) {
            const bf16x8 *Fw = L == 0 ? F2 : F3;
            const float *bl = bias + 128 * (L + 1);
            f32x16 a[2][2];
#pragma unroll
            for (int np = 0; np < 2; ++np) {
                a[np][0] = f32x16{};
                a[np][1] = f32x16{};
#pragma unroll
                for (int kc = 0; kc < 8; ++kc) {
                    const bf16x8 w0 = Fw[((2 * np) * 8 + kc) * 64 + l], w1 = Fw[((2 * np + 1) * 8 + kc) * 64 + l];
                    a[np][0] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(w0, h[kc], a[np][0], 0, 0, 0);
                    a[np][1] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(w1, h[kc], a[np][1], 0, 0, 0);
                }
            }
#pragma unroll
            for (int np = 0; np < 2; ++np)
#pragma unroll
                for (int r = 0; r < 16; ++r) {
                    const int n0 = (2 * np) * 32 + 8 * (r >> 2) + 4 * half + (r & 3);
                    h[4 * np + (r >> 3)][r & 7] = f2bf(softplus2_f(a[np][0][r] + bl[n0]));
                    h[4 * np + 2 + (r >> 3)][r & 7] = f2bf(softplus2_f(a[np][1][r] + bl[n0 + 32]));
                }
            __builtin_amdgcn_sched_group_barrier(0x008, 16, 0);     // MFMA
#pragma unroll
            for (int i = 0; i < 16; ++i) {
                __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);  // MFMA
                __builtin_amdgcn_sched_group_barrier(0x002, 12, 0); // VALU
            }
        }
        // layer 4 (linear), fused output transform:
        // y = acc ystd + (b4 ystd + ymean); tau = e^y col_dry, pfrac = max(y, 0).
        // The operands are swapped (rows x K times K x features): the layer-3
        // output already sits in the A-operand layout, and the weight fragments
        // in the B-operand layout, so the result has one feature per lane and
        // 16 rows per lane, and each store writes one 128-byte line per half
        // wavefront (32 consecutive features of one row) instead of 32 partial
        // lines.  A lane's 16 rows (output row, dry-air column) come from the
        // lanes that prepared them.  Rows past the end of the launch (a last
        // tile that is not full) were computed from the last row's inputs, so
        // they store that row's values again, bit for bit: every store is
        // unconditional, with no branch per store.  The stores are streaming
        // (nontemporal): the 8.5 GB a 65,536-column launch writes are read
        // back by the solver, not by this kernel, and do not displace the
        // inputs in the caches (2.58 -> 2.15 ms per launch).
        int orr[16];
        float cdr[16];
        {
            const int mine = (int)orow;
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int row = 8 * (r >> 2) + 4 * half + (r & 3);
                orr[r] = __shfl(mine, row);
                cdr[r] = __shfl(col_dry, row);
            }
        }
        // output tiles 2 np, 2 np + 1 (tau for np < 2, Planck fraction after):
        // the products of one pair run while the previous pair's outputs are
        // transformed and stored
        auto chain = [&](int np, f32x16 (&acc)[2]) {
            acc[0] = f32x16{};
            acc[1] = f32x16{};
#pragma unroll
            for (int kc = 0; kc < 8; ++kc) {
                const bf16x8 w0 = F4[((2 * np) * 8 + kc) * 64 + l], w1 = F4[((2 * np + 1) * 8 + kc) * 64 + l];
                acc[0] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(h[kc], w0, acc[0], 0, 0, 0);
                acc[1] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(h[kc], w1, acc[1], 0, 0, 0);
            }
        };
        auto store = [&](int np, const f32x16 (&acc)[2]) {
            const bool tau = np < 2;
            float *out_p = tau ? A.tau : A.pfrac;
#pragma unroll
            for (int u = 0; u < 2; ++u) {
                const int feat = (2 * np + u) * 32 + rl;
                const float sc = ysd[feat], cc = yc[feat];
                const int col = feat - (tau ? 0 : MLP_G);
#pragma unroll
                for (int r = 0; r < 16; ++r) {
                    const float y = acc[u][r] * sc + cc;
                    const float v = tau ? __expf(y) * cdr[r] : fmaxf(y, 0.0f);
                    __builtin_nontemporal_store(v, &out_p[(long)orr[r] * MLP_G + col]);
                }
            }
        };
        f32x16 acc[2][2];
        chain(0, acc[0]);
#pragma unroll
        for (int np = 1; np < 4; ++np) {
            chain(np, acc[np & 1]);
            store(np - 1, acc[(np - 1) & 1]);
            __builtin_amdgcn_sched_group_barrier(0x008, 2, 1);      // MFMA
#pragma unroll
            for (int i = 0; i < 14; ++i) {
                __builtin_amdgcn_sched_group_barrier(0x008, 1, 1);  // MFMA
                __builtin_amdgcn_sched_group_barrier(0x002, 6, 1);  // VALU
                __builtin_amdgcn_sched_group_barrier(0x040, 2, 1);  // VMEM write
            }
        }
        store(3, acc[1]);
        if (valid && half == 0) {
            A.oob[orow] = oob ? 1 : 0;
            if (oob) A.colflag[dcol] = 1;
        }
    }
}

}  // namespace ccpprad

// ---------------------------------------------------------------------------
// C-ABI of the emulator
// ---------------------------------------------------------------------------
#include <vector>
#include <string.h>
#include <stdio.h>
#include <math.h>
#include "../../include/ccpprad.h"

using namespace ccpprad;

struct ccpprad_mlp {
    float *p = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;   // FP32 master, grads, Adam
    __bf16 *p16 = nullptr;                                          // BF16 copies
    unsigned char *image = nullptr;                                 // inference LDS image
    float *ymean = nullptr, *ystd = nullptr;
    float hymean[MLP_OUT] = {}, hystd[MLP_OUT] = {};   // host copies for the inference image
    float xmean[MLP_IN], xistd[MLP_IN], lo[MLP_IN], hi[MLP_IN];
    int step = 0;
    // training work buffers (rows capacity)
    int cap = 0;
    float *z[4] = {nullptr, nullptr, nullptr, nullptr};   // pre-activations
    __bf16 *h16[4] = {nullptr, nullptr, nullptr, nullptr};  // activations (h16[0] = input copy slot unused)
    float *d32 = nullptr, *dpart = nullptr, *lpart = nullptr;
    __bf16 *d16a = nullptr, *d16b = nullptr;
    float *loss = nullptr;
    __bf16 *x14 = nullptr, *w1e = nullptr;   // layer-1 input (hi/lo) and [W1; W1]
    float *g14 = nullptr;                    // dW1eff
};

static const int LSZ[5] = {MLP_IN, MLP_H, MLP_H, MLP_H, MLP_OUT};
// offsets of W_i (K x N, input-major) and b_i in the flat parameter vector
static long woff(int i) { long o = 0; for (int k = 0; k < i; ++k) o += (long)LSZ[k] * LSZ[k + 1] + LSZ[k + 1]; return o; }
static long boff(int i) { return woff(i) + (long)LSZ[i] * LSZ[i + 1]; }

static int mlp_seterr(char *e, int n, const char *m)
{
    if (e && n > 0) snprintf(e, n, "%s", m);
    return 1;
}

extern "C" int ccpprad_mlp_create(ccpprad_mlp **out, char *errmsg, int errlen)
{
    if (!out) return mlp_seterr(errmsg, errlen, "null output pointer");
    ccpprad_mlp *m = new ccpprad_mlp();
    hipError_t e = hipSuccess;
    e = (hipError_t)(e | hipMalloc(&m->p, sizeof(float) * MLP_NPARAM));
    e = (hipError_t)(e | hipMalloc(&m->g, sizeof(float) * MLP_NPARAM));
    e = (hipError_t)(e | hipMalloc(&m->m, sizeof(float) * MLP_NPARAM));
    e = (hipError_t)(e | hipMalloc(&m->v, sizeof(float) * MLP_NPARAM));
    e = (hipError_t)(e | hipMalloc(&m->p16, sizeof(__bf16) * MLP_NPARAM));
    e = (hipError_t)(e | hipMalloc(&m->image, MLP_LDS_BYTES));
    e = (hipError_t)(e | hipMalloc(&m->ymean, sizeof(float) * MLP_OUT));
    e = (hipError_t)(e | hipMalloc(&m->ystd, sizeof(float) * MLP_OUT));
    e = (hipError_t)(e | hipMalloc(&m->loss, sizeof(float) * 2));
    e = (hipError_t)(e | hipMalloc(&m->w1e, sizeof(__bf16) * 14 * 128));
    e = (hipError_t)(e | hipMalloc(&m->g14, sizeof(float) * 14 * 128));
    if (e != hipSuccess) { delete m; return mlp_seterr(errmsg, errlen, "device allocation failed"); }
    (void)hipMemset(m->m, 0, sizeof(float) * MLP_NPARAM);
    (void)hipMemset(m->v, 0, sizeof(float) * MLP_NPARAM);
    for (int i = 0; i < MLP_IN; ++i) { m->xmean[i] = 0; m->xistd[i] = 1; m->lo[i] = -1e30f; m->hi[i] = 1e30f; }
    std::vector<float> one(MLP_OUT, 1.0f), zero(MLP_OUT, 0.0f);
    (void)hipMemcpy(m->ymean, zero.data(), sizeof(float) * MLP_OUT, hipMemcpyHostToDevice);
    (void)hipMemcpy(m->ystd, one.data(), sizeof(float) * MLP_OUT, hipMemcpyHostToDevice);
    *out = m;
    return 0;
}

extern "C" void ccpprad_mlp_destroy(ccpprad_mlp *m)
{
    if (!m) return;
    void *ps[] = {m->p, m->g, m->m, m->v, m->p16, m->image, m->ymean, m->ystd, m->loss, m->z[0], m->z[1], m->z[2],
                  m->z[3], m->h16[0], m->h16[1], m->h16[2], m->h16[3], m->d32, m->dpart, m->lpart, m->d16a, m->d16b,
                  m->x14, m->w1e, m->g14};
    for (void *p : ps) if (p) (void)hipFree(p);
    delete m;
}

// Build the inference LDS image from the FP32 parameters (host side).
static void build_image(const std::vector<float> &P, const float *ystd, const float *ymean, std::vector<unsigned char> &img)
{
    img.assign(MLP_LDS_BYTES, 0);
    __bf16 *fr = (__bf16 *)img.data();
    auto W = [&](int layer, int k, int n) -> float {
        return P[woff(layer) + (long)k * LSZ[layer + 1] + n];
    };
    int f = 0;
    // layer 1: K slots 0..15 = features (0..6), zero padded
    for (int nt = 0; nt < 4; ++nt, ++f)
        for (int l = 0; l < 64; ++l)
            for (int j = 0; j < 8; ++j) {
                const int m = nt * 32 + (l & 31), k = 8 * (l >> 5) + j;
                fr[((long)f * 64 + l) * 8 + j] =
                    (__bf16)(k < MLP_IN ? W(0, k, m) : (k < MLP_K1 ? W(0, k - MLP_IN, m) : 0.0f));
            }
    // layers 2..4: K order permuted to the previous layer's D-register order
    for (int layer = 1; layer <= 3; ++layer) {
        const int ntn = LSZ[layer + 1] / 32;
        for (int nt = 0; nt < ntn; ++nt)
            for (int kc = 0; kc < 8; ++kc, ++f)
                for (int l = 0; l < 64; ++l)
                    for (int j = 0; j < 8; ++j) {
                        const int m = nt * 32 + (l & 31), h = l >> 5;
                        const int k = 16 * kc + 8 * (j >> 2) + 4 * h + (j & 3);
                        fr[((long)f * 64 + l) * 8 + j] = (__bf16)W(layer, k, m);
                    }
    }
    float *b = (float *)(img.data() + NFRAG * FRAG_BYTES);
    // hidden-layer biases in the base-2 domain of the inference kernel
    for (int i = 0; i < 3; ++i)
        for (int n = 0; n < 128; ++n) b[128 * i + n] = (float)((double)P[boff(i) + n] * LOG2E_D);
    for (int n = 0; n < 256; ++n) b[384 + n] = P[boff(3) + n];
    // output de-standardisation folded with the layer-4 bias:
    // y = (acc + b4) ystd + ymean = acc ystd + (b4 ystd + ymean), where the
    // kernel's acc is the natural one over ln2 (base-2 hidden activations)
    for (int n = 0; n < 256; ++n) {
        b[640 + n] = (float)((double)ystd[n] * LN2_D);
        b[896 + n] = (float)((double)P[boff(3) + n] * ystd[n] + ymean[n]);
    }
}

static int refresh(ccpprad_mlp *m)
{
    std::vector<float> P(MLP_NPARAM);
    if (hipMemcpy(P.data(), m->p, sizeof(float) * MLP_NPARAM, hipMemcpyDeviceToHost) != hipSuccess) return 2;
    std::vector<unsigned char> img;
    build_image(P, m->hystd, m->hymean, img);
    if (hipMemcpy(m->image, img.data(), img.size(), hipMemcpyHostToDevice) != hipSuccess) return 2;
    hipLaunchKernelGGL(to_bf16_kernel, dim3((MLP_NPARAM + 255) / 256), dim3(256), 0, 0, m->p, m->p16, (long)MLP_NPARAM);
    return hipDeviceSynchronize() == hipSuccess ? 0 : 2;
}

extern "C" int ccpprad_mlp_nparams(void) { return MLP_NPARAM; }

extern "C" int ccpprad_mlp_set_params(ccpprad_mlp *m, const float *host_params, char *errmsg, int errlen)
{
    if (!m || !host_params) return mlp_seterr(errmsg, errlen, "null argument");
    if (hipMemcpy(m->p, host_params, sizeof(float) * MLP_NPARAM, hipMemcpyHostToDevice) != hipSuccess)
        return mlp_seterr(errmsg, errlen, "copy failed");
    return refresh(m) ? mlp_seterr(errmsg, errlen, "refresh failed") : 0;
}

extern "C" int ccpprad_mlp_get_params(ccpprad_mlp *m, float *host_params, float *host_grads)
{
    if (!m) return 1;
    if (host_params && hipMemcpy(host_params, m->p, sizeof(float) * MLP_NPARAM, hipMemcpyDeviceToHost) != hipSuccess) return 2;
    if (host_grads && hipMemcpy(host_grads, m->g, sizeof(float) * MLP_NPARAM, hipMemcpyDeviceToHost) != hipSuccess) return 2;
    return 0;
}

extern "C" int ccpprad_mlp_set_norm(ccpprad_mlp *m, const float *xmean, const float *xstd, const float *lo,
                                    const float *hi, const float *ymean, const float *ystd, char *errmsg, int errlen)
{
    if (!m || !xmean || !xstd || !lo || !hi || !ymean || !ystd) return mlp_seterr(errmsg, errlen, "null argument");
    for (int i = 0; i < MLP_IN; ++i) {
        if (!(xstd[i] > 0)) return mlp_seterr(errmsg, errlen, "feature std must be positive");
        m->xmean[i] = xmean[i]; m->xistd[i] = 1.0f / xstd[i]; m->lo[i] = lo[i]; m->hi[i] = hi[i];
    }
    if (hipMemcpy(m->ymean, ymean, sizeof(float) * MLP_OUT, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(m->ystd, ystd, sizeof(float) * MLP_OUT, hipMemcpyHostToDevice) != hipSuccess)
        return mlp_seterr(errmsg, errlen, "copy failed");
    for (int n = 0; n < MLP_OUT; ++n) { m->hymean[n] = ymean[n]; m->hystd[n] = ystd[n]; }
    return refresh(m) ? mlp_seterr(errmsg, errlen, "refresh failed") : 0;
}

static int ensure_cap(ccpprad_mlp *m, int rows)
{
    if (rows <= m->cap) return 0;
    void *ps[] = {m->z[0], m->z[1], m->z[2], m->z[3], m->h16[0], m->h16[1], m->h16[2], m->h16[3], m->d32, m->dpart,
                  m->lpart, m->d16a, m->d16b, m->x14};
    for (void *p : ps) if (p) (void)hipFree(p);
    hipError_t e = hipSuccess;
    for (int i = 0; i < 4; ++i) {
        e = (hipError_t)(e | hipMalloc(&m->z[i], sizeof(float) * (size_t)rows * LSZ[i + 1]));
        e = (hipError_t)(e | hipMalloc(&m->h16[i], sizeof(__bf16) * (size_t)rows * LSZ[i + 1]));
    }
    e = (hipError_t)(e | hipMalloc(&m->d32, sizeof(float) * (size_t)rows * MLP_OUT));
    e = (hipError_t)(e | hipMalloc(&m->d16a, sizeof(__bf16) * (size_t)rows * MLP_OUT));
    e = (hipError_t)(e | hipMalloc(&m->d16b, sizeof(__bf16) * (size_t)rows * MLP_OUT));
    e = (hipError_t)(e | hipMalloc(&m->dpart, sizeof(float) * (size_t)64 * MLP_H * MLP_OUT));
    e = (hipError_t)(e | hipMalloc(&m->lpart, sizeof(float) * (size_t)((rows * (size_t)MLP_OUT + 255) / 256 + 1)));
    e = (hipError_t)(e | hipMalloc(&m->x14, sizeof(__bf16) * (size_t)rows * 14));
    if (e != hipSuccess) { m->cap = 0; return 2; }
    m->cap = rows;
    return 0;
}

// Forward pass of a batch: X (rows x 7, BF16 standardised features) -> Y
// (rows x 256, FP32, standardised outputs).  Keeps pre-activations for backward.
static void forward(ccpprad_mlp *m, const float *X, int rows, float *Y, hipStream_t st)
{
    hipLaunchKernelGGL(prep_x14_kernel, dim3((unsigned)(((long)rows * MLP_IN + 255) / 256)), dim3(256), 0, st, X, rows, m->x14);
    hipLaunchKernelGGL(w1eff_kernel, dim3((14 * 128 + 255) / 256), dim3(256), 0, st, m->p16, m->w1e);
    const __bf16 *in = m->x14;
    for (int i = 0; i < 4; ++i) {
        const int K = i == 0 ? MLP_K1 : LSZ[i];
        GemmArgs g{};
        g.M = rows; g.N = LSZ[i + 1]; g.K = K;
        g.A = in; g.sam = K; g.sak = 1;
        g.B = i == 0 ? m->w1e : m->p16 + woff(i); g.sbk = LSZ[i + 1]; g.sbn = 1;
        g.bias = m->p + boff(i);
        g.C32 = (i == 3 && Y) ? Y : m->z[i]; g.ldc = LSZ[i + 1];
        g.C16 = i < 3 ? m->h16[i] : nullptr; g.ldc16 = LSZ[i + 1];
        g.act = i < 3;
        gemm(g, 0, st);
        in = m->h16[i];
    }
}

extern "C" int ccpprad_mlp_forward_dev(ccpprad_mlp *m, const float *X, int rows, float *Y, void *stream,
                                       char *errmsg, int errlen)
{
    if (!m || !X || !Y || rows <= 0) return mlp_seterr(errmsg, errlen, "bad arguments");
    if (ensure_cap(m, rows)) return mlp_seterr(errmsg, errlen, "device allocation failed");
    forward(m, X, rows, Y, (hipStream_t)stream);
    return hipGetLastError() == hipSuccess ? 0 : mlp_seterr(errmsg, errlen, "launch failed");
}

// Loss gradients of a batch into m->g (no update).  T: rows x 256 FP32.
static void backward(ccpprad_mlp *m, const float *X, const float *T, int rows, hipStream_t st)
{
    const long n = (long)rows * MLP_OUT;
    forward(m, X, rows, m->z[3], st);
    const int nb = (int)((n + 255) / 256);
    hipLaunchKernelGGL(mse_grad_kernel, dim3(nb), dim3(256), 0, st, m->z[3], T, n, 1.0f / (float)n, m->d32, m->d16a,
                       m->lpart);
    hipLaunchKernelGGL(sum_kernel, dim3(1), dim3(256), 0, st, m->lpart, nb, 1.0f / (float)n, m->loss);
    __bf16 *dcur = m->d16a, *dnext = m->d16b;
    for (int i = 3; i >= 0; --i) {
        const int K = i == 0 ? MLP_K1 : LSZ[i], N = LSZ[i + 1];
        const __bf16 *Hin = i == 0 ? m->x14 : m->h16[i - 1];
        // dW_i = H_in^T dZ_i (K x N), split over rows
        const int splits = std::min(64, (rows + 511) / 512);
        const int kper = ((rows + splits - 1) / splits + 31) / 32 * 32;
        GemmArgs g{};
        g.M = K; g.N = N; g.K = rows;
        g.A = Hin; g.sam = 1; g.sak = K;
        g.B = dcur; g.sbk = N; g.sbn = 1;
        g.C32 = m->dpart; g.ldc = N; g.ksplit = kper;
        const int nsplit = (rows + kper - 1) / kper;
        gemm(g, nsplit, st);
        hipLaunchKernelGGL(sum_partials_kernel, dim3((K * N + 255) / 256), dim3(256), 0, st, m->dpart, nsplit,
                           (long)K * N, i == 0 ? m->g14 : m->g + woff(i));
        if (i == 0)
            hipLaunchKernelGGL(fold_w1_kernel, dim3((7 * 128 + 255) / 256), dim3(256), 0, st, m->g14, m->g + woff(0));
        // db_i = column sums of dZ_i (FP32 copy kept in d32)
        hipLaunchKernelGGL(colsum_kernel, dim3((N + 127) / 128), dim3(128), 0, st, m->d32, rows, N, (long)N,
                           m->g + boff(i));
        if (i == 0) break;
        // dZ_{i-1} = (dZ_i W_i^T) * sigmoid(Z_{i-1})
        GemmArgs h{};
        h.M = rows; h.N = K; h.K = N;
        h.A = dcur; h.sam = N; h.sak = 1;
        h.B = m->p16 + woff(i); h.sbk = 1; h.sbn = N;     // W^T(k=n_out, n=k_in)
        h.Z = m->z[i - 1]; h.ldz = K;
        h.C32 = m->d32; h.ldc = K;
        h.C16 = dnext; h.ldc16 = K;
        gemm(h, 0, st);
        std::swap(dcur, dnext);
    }
}

extern "C" int ccpprad_mlp_grads_dev(ccpprad_mlp *m, const float *X, const float *T, int rows, float *loss_out,
                                     void *stream, char *errmsg, int errlen)
{
    if (!m || !X || !T || rows <= 0) return mlp_seterr(errmsg, errlen, "bad arguments");
    if (ensure_cap(m, rows)) return mlp_seterr(errmsg, errlen, "device allocation failed");
    hipStream_t st = (hipStream_t)stream;
    backward(m, X, T, rows, st);
    if (loss_out) (void)hipMemcpyAsync(loss_out, m->loss, sizeof(float), hipMemcpyDeviceToDevice, st);
    return hipGetLastError() == hipSuccess ? 0 : mlp_seterr(errmsg, errlen, "launch failed");
}

extern "C" int ccpprad_mlp_train_step_dev(ccpprad_mlp *m, const float *X, const float *T, int rows, float lr,
                                          float *loss_out, void *stream, char *errmsg, int errlen)
{
    int rc = ccpprad_mlp_grads_dev(m, X, T, rows, loss_out, stream, errmsg, errlen);
    if (rc) return rc;
    m->step += 1;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.0f - powf(b1, (float)m->step), c2 = 1.0f - powf(b2, (float)m->step);
    hipLaunchKernelGGL(adam_kernel, dim3((MLP_NPARAM + 255) / 256), dim3(256), 0, (hipStream_t)stream, m->p, m->g,
                       m->m, m->v, m->p16, MLP_NPARAM, lr, b1, b2, 1e-8f, c1, c2);
    return hipGetLastError() == hipSuccess ? 0 : mlp_seterr(errmsg, errlen, "launch failed");
}

// Rebuild the inference weights after training (host round trip of 268 KB).
extern "C" int ccpprad_mlp_finalize_training(ccpprad_mlp *m, char *errmsg, int errlen)
{
    if (!m) return mlp_seterr(errmsg, errlen, "null argument");
    return refresh(m) ? mlp_seterr(errmsg, errlen, "refresh failed") : 0;
}

namespace ccpprad {
void launch_mlp_infer(const ::ccpprad_mlp *m, MlpInferArgs &a, hipStream_t st)
{
    a.image = m->image;
    a.ymean = m->ymean;
    a.ystd = m->ystd;
    for (int i = 0; i < MLP_NIN; ++i) { a.xmean[i] = m->xmean[i]; a.xistd[i] = (float)((double)m->xistd[i] * LOG2E_D); a.lo[i] = m->lo[i]; a.hi[i] = m->hi[i]; }
    static bool attr = false;
    if (!attr) {
        (void)hipFuncSetAttribute((const void *)mlp_infer_kernel, hipFuncAttributeMaxDynamicSharedMemorySize, MLP_LDS_BYTES);
        attr = true;
    }
    const long rows = (long)a.ncols * a.nlay;
    const long tiles = (rows + 31) / 32;
    long grid = (tiles + 7) / 8;
    if (grid > 256L * 8) grid = 256L * 8;
    if (grid < 1) grid = 1;
    hipLaunchKernelGGL(mlp_infer_kernel, dim3((unsigned)grid), dim3(512), MLP_LDS_BYTES, st, a);
}
}  // namespace ccpprad

// Stand-alone inference entry (tests / emulator evaluation): rows of
// (ncols x nlay) from CCPP-layout inputs -> tau, pfrac [d][l][g] and the
// out-of-hull flags.
extern "C" int ccpprad_mlp_infer_dev(ccpprad_mlp *m, int nCol, int nLay, const double *p_lay, const double *p_lev,
                                     const double *t_lay, const double *vmr_h2o, const double *vmr_o3,
                                     const double *vmr_co2, const double *vmr_ch4, const double *vmr_n2o, float *tau,
                                     float *pfrac, unsigned char *oob, unsigned char *colflag, void *stream,
                                     char *errmsg, int errlen)
{
    if (!m || !p_lay || !p_lev || !t_lay || !vmr_h2o || !vmr_o3 || !vmr_co2 || !vmr_ch4 || !vmr_n2o || !tau ||
        !pfrac || !oob || !colflag || nCol <= 0 || nLay <= 0)
        return mlp_seterr(errmsg, errlen, "bad arguments");
    MlpInferArgs a{};
    a.nlay = nLay; a.ncol_ld = nCol; a.col0 = 0; a.ncols = nCol;
    a.p_lay = p_lay; a.p_lev = p_lev; a.t_lay = t_lay;
    a.vmr[0] = vmr_h2o; a.vmr[1] = vmr_o3; a.vmr[2] = vmr_co2; a.vmr[3] = vmr_ch4; a.vmr[4] = vmr_n2o;
    a.tau = tau; a.pfrac = pfrac; a.oob = oob; a.colflag = colflag;
    hipStream_t st = (hipStream_t)stream;
    launch_fill_u8(colflag, (size_t)nCol, st);
    launch_mlp_infer(m, a, st);
    return hipGetLastError() == hipSuccess ? 0 : mlp_seterr(errmsg, errlen, "launch failed");
}
