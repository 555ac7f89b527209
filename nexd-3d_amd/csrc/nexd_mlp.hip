// libnexd_gpu.so, the nexd_mlp_* part of include/nexd_gpu.h: a multilayer
// perceptron (GELU hidden layers, linear output) trained and evaluated by
// this file's kernels.  Weights are kept in FP32 (the Adam master copy) and
// in BF16 (the GEMM operand); every product runs on the matrix cores as
// v_mfma_f32_16x16x32_bf16 with FP32 accumulation.
//
// Layer l maps in_l -> out_l:  Z = A_prev W^T + b,  A = GELU(Z) (hidden),
// Y = Z (output).  Shapes on the device are padded to multiples of 64
// (rows of the batch, every feature dimension), with zero padding that
// stays zero through training (its inputs and gradients are zero).
//
// No counterpart in the reference: this is the learned source-to-
// seismogram surrogate that is trained on, and judged by, the solver.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/nexd_gpu.h"

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef uint16_t bf16_t;   // storage of one bfloat16

// round to nearest even (inputs are finite)
__host__ __device__ inline bf16_t to_bf16(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (bf16_t)(u >> 16);
}

// GELU with the error function, and its derivative
__device__ inline float gelu(float x) { return 0.5f * x * (1.0f + erff(x * 0.70710678118654752f)); }
__device__ inline float gelu_grad(float x) {
    const float cdf = 0.5f * (1.0f + erff(x * 0.70710678118654752f));
    const float pdf = 0.39894228040143268f * expf(-0.5f * x * x);
    return cdf + x * pdf;
}

// ---------------------------------------------------------------------------
// GEMM  C[M][N] = sum_k A(m,k) B(k,n)   (M, N multiples of 64, K of 32)
//   A_KC: A stored [M][K] (k contiguous), else [K][M] (m contiguous)
//   B_KC: B stored [N][K] (k contiguous), else [K][N] (n contiguous)
// Block 256 threads = 4 waves in 2 x 2, block tile 64 x 64, each wave
// 32 x 32 = 2 x 2 MFMA tiles of 16 x 16, k-tiles of 32 staged in LDS as
// [row][k] with a padded row (40 bf16) so that the 16-byte fragment reads
// of 16 rows fall in different banks.  The next k-tile's global loads are
// issued before the current one's products.
// ---------------------------------------------------------------------------
constexpr int BM = 64, BN = 64, BK = 32, LROW = BK + 8;

enum Epi {
    EPI_STORE = 0,        // out_f32 = acc                         (weight gradients)
    EPI_BIAS_GELU = 1,    // out_f32 = acc + b; out_bf16 = GELU    (hidden layer forward)
    EPI_BIAS = 2,         // out_f32 = acc + b                     (output layer forward)
    EPI_GELU_GRAD = 3,    // g = acc * GELU'(aux); out_f32 = g, out_bf16 = g   (backward)
};

struct GemmArgs {
    const bf16_t *A, *B;
    int M, N, K, lda, ldb;
    float *out_f32;
    bf16_t *out_bf16;
    int ldc;
    const float *bias;    // [N]
    const float *aux;     // [M][ldc] pre-activations for EPI_GELU_GRAD
};

// one thread's part of a 64 x 32 tile: 8 consecutive elements (16 bytes)
struct Frag {
    uint4 v;
};

template <bool KC>
__device__ inline Frag load_tile(const bf16_t *P, int ld, int row0, int k0, int t) {
    // KC: rows row0..row0+63 of [rows][ld], k0..k0+31: row t/4, k chunk (t%4)*8
    // else: [K][rows]: k row t/8 (0..31), row chunk (t%8)*8
    Frag f;
    if constexpr (KC) f.v = *reinterpret_cast<const uint4 *>(P + (size_t)(row0 + t / 4) * ld + k0 + (t % 4) * 8);
    else f.v = *reinterpret_cast<const uint4 *>(P + (size_t)(k0 + t / 8) * ld + row0 + (t % 8) * 8);
    return f;
}

template <bool KC>
__device__ inline void store_tile(bf16_t *S, const Frag &f, int t) {
    if constexpr (KC) {
        *reinterpret_cast<uint4 *>(S + (t / 4) * LROW + (t % 4) * 8) = f.v;
    } else {
        // transpose through LDS: 8 rows of one k
        const bf16_t *h = reinterpret_cast<const bf16_t *>(&f.v);
        const int k = t / 8, r0 = (t % 8) * 8;
#pragma unroll
        for (int i = 0; i < 8; ++i) S[(r0 + i) * LROW + k] = h[i];
    }
}

template <bool A_KC, bool B_KC, int EPI>
__global__ void __launch_bounds__(256) gemm_kernel(GemmArgs g) {
    __shared__ __align__(16) bf16_t sA[2][BM * LROW];
    __shared__ __align__(16) bf16_t sB[2][BN * LROW];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int nbn = g.N / BN;
    const int m0 = (blockIdx.x / nbn) * BM, n0 = (blockIdx.x % nbn) * BN;
    const int wm = (w >> 1) * 32, wn = (w & 1) * 32;
    f32x4 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};

    const int nk = g.K / BK;
    Frag fa = load_tile<A_KC>(g.A, g.lda, m0, 0, t);
    Frag fb = load_tile<B_KC>(g.B, g.ldb, n0, 0, t);
    store_tile<A_KC>(sA[0], fa, t);
    store_tile<B_KC>(sB[0], fb, t);
    __syncthreads();
    for (int kt = 0; kt < nk; ++kt) {
        const int cur = kt & 1;
        if (kt + 1 < nk) {
            fa = load_tile<A_KC>(g.A, g.lda, m0, (kt + 1) * BK, t);
            fb = load_tile<B_KC>(g.B, g.ldb, n0, (kt + 1) * BK, t);
        }
        // fragments: lane l holds row l%16 and k 8*(l/16) .. +8 of its tile
        bf16x8 a[2], b[2];
#pragma unroll
        for (int i = 0; i < 2; ++i)
            a[i] = *reinterpret_cast<const bf16x8 *>(sA[cur] + (wm + i * 16 + (lane & 15)) * LROW + 8 * (lane >> 4));
#pragma unroll
        for (int j = 0; j < 2; ++j)
            b[j] = *reinterpret_cast<const bf16x8 *>(sB[cur] + (wn + j * 16 + (lane & 15)) * LROW + 8 * (lane >> 4));
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
        if (kt + 1 < nk) {
            store_tile<A_KC>(sA[cur ^ 1], fa, t);
            store_tile<B_KC>(sB[cur ^ 1], fb, t);
        }
        __syncthreads();
    }
    // accumulator element v of tile (i, j): row 4*(l/16)+v, column l%16
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int m = m0 + wm + i * 16 + 4 * (lane >> 4) + v;
                const int n = n0 + wn + j * 16 + (lane & 15);
                const size_t o = (size_t)m * g.ldc + n;
                float x = acc[i][j][v];
                if constexpr (EPI == EPI_STORE) {
                    g.out_f32[o] = x;
                } else if constexpr (EPI == EPI_BIAS_GELU) {
                    x += g.bias[n];
                    g.out_f32[o] = x;
                    g.out_bf16[o] = to_bf16(gelu(x));
                } else if constexpr (EPI == EPI_BIAS) {
                    g.out_f32[o] = x + g.bias[n];
                } else {
                    x *= gelu_grad(g.aux[o]);
                    g.out_f32[o] = x;
                    g.out_bf16[o] = to_bf16(x);
                }
            }
}

// ---------------------------------------------------------------------------
// element-wise and reduction kernels
// ---------------------------------------------------------------------------
// rows of an FP32 batch [rows][cols] (host layout) into a padded BF16 [rows][ldp]
__global__ void pack_rows_kernel(const float *src, int rows, int cols, bf16_t *dst, int ldp) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)rows * ldp) return;
    const int r = (int)(i / ldp), c = (int)(i % ldp);
    dst[i] = c < cols ? to_bf16(src[(size_t)r * cols + c]) : (bf16_t)0;
}

// gathers batch rows idx[0..B) of a dataset [n][cols] (device) into padded BF16
__global__ void gather_rows_kernel(const float *src, const int *idx, int B, int cols, bf16_t *dst, int ldp) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)B * ldp) return;
    const int r = (int)(i / ldp), c = (int)(i % ldp);
    dst[i] = c < cols ? to_bf16(src[(size_t)idx[r] * cols + c]) : (bf16_t)0;
}

// squared-error loss of the batch: per-block partial sums of (y - t)^2 and
// the gradient d = 2 (y - t) / (B * cols) in FP32 and BF16; target rows are
// idx[r] of the dataset (idx null: row r of tgt)
__global__ void __launch_bounds__(256) mse_kernel(const float *y, int B, int cols, const float *tgt, const int *idx,
                                                   float *dy_f32, bf16_t *dy_bf16, float *partial) {
    __shared__ float red[256];
    const size_t n = (size_t)B * cols;
    const float scale = 2.0f / (float)n;
    float s = 0.f;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const int r = (int)(i / cols), c = (int)(i % cols);
        const float d = y[i] - tgt[(size_t)(idx ? idx[r] : r) * cols + c];
        s += d * d;
        if (dy_f32) {
            dy_f32[i] = scale * d;
            dy_bf16[i] = to_bf16(scale * d);
        }
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if ((int)threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// fixed-order sum of the partials: loss = sum / (B * cols)
__global__ void __launch_bounds__(256) sum_kernel(const float *partial, int n, float inv, float *out) {
    __shared__ float red[256];
    float s = 0.f;
    for (int i = threadIdx.x; i < n; i += 256) s += partial[i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if ((int)threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = red[0] * inv;
}

// bias gradient: column sums of dz [B][ld] over the batch
__global__ void colsum_kernel(const float *dz, int B, int cols, int ld, float *db) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= cols) return;
    float s = 0.f;
    for (int r = 0; r < B; ++r) s += dz[(size_t)r * ld + c];
    db[c] = s;
}

// Adam (Kingma and Ba, 2015) with bias correction folded into the step:
// m = b1 m + (1-b1) g, v = b2 v + (1-b2) g^2, w -= lr_t m / (sqrt(v) + eps_t);
// the BF16 copy of w follows.  Padding entries (mask 0) stay zero.
__global__ void adam_kernel(float *w, bf16_t *wb, const float *grad, float *m, float *v, size_t n, float lr_t, float b1,
                            float b2, float eps_t, const unsigned char *mask) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (mask && !mask[i]) return;
    const float gi = grad[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float wi = w[i] - lr_t * mi / (sqrtf(vi) + eps_t);
    w[i] = wi;
    if (wb) wb[i] = to_bf16(wi);
}

__global__ void to_bf16_kernel(const float *src, bf16_t *dst, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = to_bf16(src[i]);
}

int pad64(int x) { return (x + 63) / 64 * 64; }

}  // namespace

// ===========================================================================
// the network object
// ===========================================================================
struct nexd_mlp {
    int nl = 0;                       // layers
    std::vector<int> in, out;         // true widths
    std::vector<int> inp, outp;       // padded widths
    int max_batch = 0, device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::string err;
    std::vector<void *> allocs;
    // per layer
    std::vector<float *> W, b, gW, gb, mW, vW, mb, vb;
    std::vector<bf16_t *> Wb;
    std::vector<unsigned char *> maskW;
    // activations: act[0] = input (bf16), act[l+1] = GELU(Z_l) (bf16), Z[l] f32
    std::vector<bf16_t *> act;
    std::vector<float *> Z;
    std::vector<float *> dz;          // f32 gradient w.r.t. Z_l
    std::vector<bf16_t *> dzb;        // bf16 copy (GEMM operand)
    float *partial = nullptr, *loss = nullptr;
    float *xin = nullptr;             // [max_batch][in_0] FP32 inputs of nexd_mlp_forward
    // dataset
    float *dx = nullptr, *dt = nullptr;
    int dn = 0;
    int *idx = nullptr;
    int adam_t = 0;
    float last_ms = 0.f;
};

namespace {

int mfail(nexd_mlp *m, int code, const std::string &s) {
    m->err = s;
    return code;
}

#define MCHK(m, expr)                                                                         \
    do {                                                                                      \
        hipError_t _e = (expr);                                                               \
        if (_e != hipSuccess) return mfail(m, NEXD_E_HIP, std::string(#expr ": ") + hipGetErrorString(_e)); \
    } while (0)

template <typename T>
int malloc_zero(nexd_mlp *m, T **p, size_t n) {
    void *v = nullptr;
    MCHK(m, hipMalloc(&v, (n ? n : 1) * sizeof(T)));
    m->allocs.push_back(v);
    MCHK(m, hipMemset(v, 0, (n ? n : 1) * sizeof(T)));
    *p = static_cast<T *>(v);
    return 0;
}

template <bool A_KC, bool B_KC, int EPI>
void gemm(nexd_mlp *m, const GemmArgs &g) {
    const int blocks = (g.M / BM) * (g.N / BN);
    hipLaunchKernelGGL((gemm_kernel<A_KC, B_KC, EPI>), dim3(blocks), dim3(256), 0, m->stream, g);
}

unsigned nblk(size_t n, unsigned b = 256) { return (unsigned)((n + b - 1) / b); }

// forward pass of a batch already in act[0] (B rows)
void forward(nexd_mlp *m, int B) {
    for (int l = 0; l < m->nl; ++l) {
        GemmArgs g{};
        g.A = m->act[l]; g.lda = m->inp[l];
        g.B = m->Wb[l]; g.ldb = m->inp[l];
        g.M = B; g.N = m->outp[l]; g.K = m->inp[l];
        g.out_f32 = m->Z[l]; g.ldc = m->outp[l]; g.bias = m->b[l];
        if (l + 1 < m->nl) {
            g.out_bf16 = m->act[l + 1];
            gemm<true, true, EPI_BIAS_GELU>(m, g);
        } else {
            gemm<true, true, EPI_BIAS>(m, g);
        }
    }
}

// backward pass from dz[nl-1] (already holding dLoss/dY): weight and bias
// gradients of every layer, dz of the hidden layers
void backward(nexd_mlp *m, int B) {
    for (int l = m->nl - 1; l >= 0; --l) {
        GemmArgs g{};
        // gW_l [outp][inp] = dz_l^T [outp x B] . act_l [B x inp]
        g.A = m->dzb[l]; g.lda = m->outp[l];         // [B][outp]: m contiguous
        g.B = m->act[l]; g.ldb = m->inp[l];          // [B][inp]: n contiguous
        g.M = m->outp[l]; g.N = m->inp[l]; g.K = B;
        g.out_f32 = m->gW[l]; g.ldc = m->inp[l];
        gemm<false, false, EPI_STORE>(m, g);
        hipLaunchKernelGGL(colsum_kernel, dim3(nblk(m->outp[l])), dim3(256), 0, m->stream, m->dz[l], B, m->outp[l],
                           m->outp[l], m->gb[l]);
        if (l == 0) break;
        // dz_{l-1} = (dz_l . W_l) * GELU'(Z_{l-1})
        GemmArgs h{};
        h.A = m->dzb[l]; h.lda = m->outp[l];         // [B][outp]: k contiguous
        h.B = m->Wb[l]; h.ldb = m->inp[l];           // W [outp][inp] as [K][N]: n contiguous
        h.M = B; h.N = m->inp[l]; h.K = m->outp[l];
        h.out_f32 = m->dz[l - 1]; h.out_bf16 = m->dzb[l - 1]; h.ldc = m->inp[l];
        h.aux = m->Z[l - 1];
        gemm<true, false, EPI_GELU_GRAD>(m, h);
    }
}

void adam(nexd_mlp *m, float lr, float b1, float b2, float eps) {
    m->adam_t += 1;
    const double c1 = 1.0 - std::pow((double)b1, m->adam_t), c2 = 1.0 - std::pow((double)b2, m->adam_t);
    const float lr_t = (float)(lr * std::sqrt(c2) / c1);
    const float eps_t = (float)(eps * std::sqrt(c2));
    for (int l = 0; l < m->nl; ++l) {
        const size_t nw = (size_t)m->outp[l] * m->inp[l];
        hipLaunchKernelGGL(adam_kernel, dim3(nblk(nw)), dim3(256), 0, m->stream, m->W[l], m->Wb[l], m->gW[l], m->mW[l],
                           m->vW[l], nw, lr_t, b1, b2, eps_t, m->maskW[l]);
        hipLaunchKernelGGL(adam_kernel, dim3(nblk(m->outp[l])), dim3(256), 0, m->stream, m->b[l], (bf16_t *)nullptr,
                           m->gb[l], m->mb[l], m->vb[l], (size_t)m->outp[l], lr_t, b1, b2, eps_t,
                           m->maskW[l] + nw);
    }
}

int check_batch(nexd_mlp *m, int B) {
    if (B <= 0 || B % 64 != 0 || B > m->max_batch)
        return mfail(m, NEXD_E_ARG, "batch must be a positive multiple of 64, at most max_batch");
    return 0;
}

}  // namespace

static std::string g_mlp_create_error;

extern "C" {

int nexd_mlp_create(int nlayers, const int *widths, int max_batch, int device, nexd_mlp **out) {
    if (!out || !widths || nlayers < 1 || nlayers > 16) { g_mlp_create_error = "bad arguments"; return NEXD_E_ARG; }
    *out = nullptr;
    if (max_batch <= 0 || max_batch % 64) { g_mlp_create_error = "max_batch must be a positive multiple of 64"; return NEXD_E_ARG; }
    for (int i = 0; i <= nlayers; ++i)
        if (widths[i] <= 0) { g_mlp_create_error = "widths must be positive"; return NEXD_E_ARG; }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) {
        g_mlp_create_error = "no such HIP device";
        return NEXD_E_HIP;
    }
    nexd_mlp *m = new nexd_mlp();
    m->nl = nlayers;
    m->max_batch = max_batch;
    m->device = device;
    auto bail = [&](int rc) { g_mlp_create_error = m->err; nexd_mlp_destroy(m); return rc; };
    if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&m->stream) != hipSuccess ||
        hipEventCreate(&m->ev0) != hipSuccess || hipEventCreate(&m->ev1) != hipSuccess)
        return bail(mfail(m, NEXD_E_HIP, "stream/event creation failed"));
    for (int l = 0; l < nlayers; ++l) {
        m->in.push_back(widths[l]);
        m->out.push_back(widths[l + 1]);
        m->inp.push_back(pad64(widths[l]));
        m->outp.push_back(pad64(widths[l + 1]));
    }
    const int L = nlayers;
    m->W.assign(L, nullptr); m->b.assign(L, nullptr); m->gW.assign(L, nullptr); m->gb.assign(L, nullptr);
    m->mW.assign(L, nullptr); m->vW.assign(L, nullptr); m->mb.assign(L, nullptr); m->vb.assign(L, nullptr);
    m->Wb.assign(L, nullptr); m->maskW.assign(L, nullptr);
    m->act.assign(L + 1, nullptr); m->Z.assign(L, nullptr); m->dz.assign(L, nullptr); m->dzb.assign(L, nullptr);
    int rc = 0;
    for (int l = 0; l < L && !rc; ++l) {
        const size_t nw = (size_t)m->outp[l] * m->inp[l], no = m->outp[l];
        (rc = malloc_zero(m, &m->W[l], nw)) || (rc = malloc_zero(m, &m->b[l], no)) ||
            (rc = malloc_zero(m, &m->gW[l], nw)) || (rc = malloc_zero(m, &m->gb[l], no)) ||
            (rc = malloc_zero(m, &m->mW[l], nw)) || (rc = malloc_zero(m, &m->vW[l], nw)) ||
            (rc = malloc_zero(m, &m->mb[l], no)) || (rc = malloc_zero(m, &m->vb[l], no)) ||
            (rc = malloc_zero(m, &m->Wb[l], nw)) || (rc = malloc_zero(m, &m->maskW[l], nw + no)) ||
            (rc = malloc_zero(m, &m->Z[l], (size_t)max_batch * no)) ||
            (rc = malloc_zero(m, &m->dz[l], (size_t)max_batch * no)) ||
            (rc = malloc_zero(m, &m->dzb[l], (size_t)max_batch * no)) ||
            (rc = malloc_zero(m, &m->act[l + 1], (size_t)max_batch * no));
        if (!rc) {
            // trainable entries: the true [out][in] block and the true biases
            std::vector<unsigned char> mk(nw + no, 0);
            for (int o = 0; o < m->out[l]; ++o) {
                for (int i = 0; i < m->in[l]; ++i) mk[(size_t)o * m->inp[l] + i] = 1;
                mk[nw + o] = 1;
            }
            if (hipMemcpy(m->maskW[l], mk.data(), mk.size(), hipMemcpyHostToDevice) != hipSuccess)
                rc = mfail(m, NEXD_E_HIP, "hipMemcpy failed");
        }
    }
    if (!rc) rc = malloc_zero(m, &m->act[0], (size_t)max_batch * m->inp[0]);
    if (!rc) (rc = malloc_zero(m, &m->partial, 1024)) || (rc = malloc_zero(m, &m->loss, 1)) ||
             (rc = malloc_zero(m, &m->idx, (size_t)max_batch)) ||
             (rc = malloc_zero(m, &m->xin, (size_t)max_batch * m->in[0]));
    if (rc) return bail(rc);
    *out = m;
    return NEXD_OK;
}

int nexd_mlp_set_params(nexd_mlp *m, int layer, const float *W, const float *b) {
    if (!m || !W || !b || layer < 0 || layer >= m->nl) return m ? mfail(m, NEXD_E_ARG, "bad layer or null") : NEXD_E_ARG;
    MCHK(m, hipSetDevice(m->device));
    const int o = m->out[layer], i = m->in[layer], ip = m->inp[layer];
    std::vector<float> wp((size_t)m->outp[layer] * ip, 0.f), bp(m->outp[layer], 0.f);
    for (int r = 0; r < o; ++r) {
        for (int c = 0; c < i; ++c) wp[(size_t)r * ip + c] = W[(size_t)r * i + c];
        bp[r] = b[r];
    }
    MCHK(m, hipMemcpy(m->W[layer], wp.data(), wp.size() * 4, hipMemcpyHostToDevice));
    MCHK(m, hipMemcpy(m->b[layer], bp.data(), bp.size() * 4, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(to_bf16_kernel, dim3(nblk(wp.size())), dim3(256), 0, m->stream, m->W[layer], m->Wb[layer],
                       wp.size());
    // a new parameter set restarts the optimiser
    MCHK(m, hipMemsetAsync(m->mW[layer], 0, wp.size() * 4, m->stream));
    MCHK(m, hipMemsetAsync(m->vW[layer], 0, wp.size() * 4, m->stream));
    MCHK(m, hipMemsetAsync(m->mb[layer], 0, bp.size() * 4, m->stream));
    MCHK(m, hipMemsetAsync(m->vb[layer], 0, bp.size() * 4, m->stream));
    m->adam_t = 0;
    MCHK(m, hipStreamSynchronize(m->stream));
    return NEXD_OK;
}

// which: 0 the FP32 weights and biases, 1 their gradients of the last
// nexd_mlp_gradients / training step
int nexd_mlp_get_params(nexd_mlp *m, int layer, int which, float *W, float *b) {
    if (!m || layer < 0 || layer >= m->nl || (which != 0 && which != 1)) return m ? mfail(m, NEXD_E_ARG, "bad argument") : NEXD_E_ARG;
    MCHK(m, hipSetDevice(m->device));
    MCHK(m, hipStreamSynchronize(m->stream));
    const int o = m->out[layer], i = m->in[layer], ip = m->inp[layer];
    std::vector<float> wp((size_t)m->outp[layer] * ip), bp(m->outp[layer]);
    MCHK(m, hipMemcpy(wp.data(), which ? m->gW[layer] : m->W[layer], wp.size() * 4, hipMemcpyDeviceToHost));
    MCHK(m, hipMemcpy(bp.data(), which ? m->gb[layer] : m->b[layer], bp.size() * 4, hipMemcpyDeviceToHost));
    for (int r = 0; r < o; ++r) {
        if (W)
            for (int c = 0; c < i; ++c) W[(size_t)r * i + c] = wp[(size_t)r * ip + c];
        if (b) b[r] = bp[r];
    }
    return NEXD_OK;
}

int nexd_mlp_forward(nexd_mlp *m, const float *x, int B, float *y) {
    if (!m || !x || !y) return m ? mfail(m, NEXD_E_ARG, "null argument") : NEXD_E_ARG;
    int rc = check_batch(m, B);
    if (rc) return rc;
    MCHK(m, hipSetDevice(m->device));
    MCHK(m, hipMemcpyAsync(m->xin, x, (size_t)B * m->in[0] * 4, hipMemcpyHostToDevice, m->stream));
    hipLaunchKernelGGL(pack_rows_kernel, dim3(nblk((size_t)B * m->inp[0])), dim3(256), 0, m->stream, m->xin, B,
                       m->in[0], m->act[0], m->inp[0]);
    forward(m, B);
    MCHK(m, hipStreamSynchronize(m->stream));
    MCHK(m, hipGetLastError());
    const int L = m->nl - 1, O = m->out[L], Op = m->outp[L];
    std::vector<float> yp((size_t)B * Op);
    MCHK(m, hipMemcpy(yp.data(), m->Z[L], yp.size() * 4, hipMemcpyDeviceToHost));
    for (int r = 0; r < B; ++r) std::memcpy(y + (size_t)r * O, yp.data() + (size_t)r * Op, (size_t)O * 4);
    return NEXD_OK;
}

// The training set [n][in] inputs and [n][out] targets, kept on the device.
int nexd_mlp_upload_dataset(nexd_mlp *m, const float *x, const float *t, int n) {
    if (!m || !x || !t || n <= 0) return m ? mfail(m, NEXD_E_ARG, "bad dataset") : NEXD_E_ARG;
    MCHK(m, hipSetDevice(m->device));
    if (m->dx) { (void)hipFree(m->dx); m->dx = nullptr; }
    if (m->dt) { (void)hipFree(m->dt); m->dt = nullptr; }
    const size_t ni = (size_t)n * m->in[0], no = (size_t)n * m->out[m->nl - 1];
    MCHK(m, hipMalloc(&m->dx, ni * 4));
    MCHK(m, hipMalloc(&m->dt, no * 4));
    MCHK(m, hipMemcpy(m->dx, x, ni * 4, hipMemcpyHostToDevice));
    MCHK(m, hipMemcpy(m->dt, t, no * 4, hipMemcpyHostToDevice));
    m->dn = n;
    return NEXD_OK;
}

// Loss and gradients of the batch of dataset rows idx[0..B) (no update);
// loss_out = mean squared error over the batch's outputs.
static int batch_gradients(nexd_mlp *m, const int *idx, int B, float *loss_out) {
    int rc = check_batch(m, B);
    if (rc) return rc;
    if (!m->dx) return mfail(m, NEXD_E_STATE, "upload a dataset first");
    for (int r = 0; r < B; ++r)
        if (idx[r] < 0 || idx[r] >= m->dn) return mfail(m, NEXD_E_ARG, "dataset row out of range");
    MCHK(m, hipMemcpyAsync(m->idx, idx, (size_t)B * 4, hipMemcpyHostToDevice, m->stream));
    hipLaunchKernelGGL(gather_rows_kernel, dim3(nblk((size_t)B * m->inp[0])), dim3(256), 0, m->stream, m->dx, m->idx, B,
                       m->in[0], m->act[0], m->inp[0]);
    forward(m, B);
    const int L = m->nl - 1;
    if (m->out[L] != m->outp[L]) return mfail(m, NEXD_E_UNSUPPORTED, "training needs an output width that is a multiple of 64");
    const int nb = 1024;
    hipLaunchKernelGGL(mse_kernel, dim3(nb), dim3(256), 0, m->stream, m->Z[L], B, m->out[L], m->dt, m->idx, m->dz[L],
                       m->dzb[L], m->partial);
    hipLaunchKernelGGL(sum_kernel, dim3(1), dim3(256), 0, m->stream, m->partial, nb,
                       1.0f / ((float)B * (float)m->out[L]), m->loss);
    backward(m, B);
    if (loss_out) {
        MCHK(m, hipMemcpyAsync(loss_out, m->loss, 4, hipMemcpyDeviceToHost, m->stream));
        MCHK(m, hipStreamSynchronize(m->stream));
    }
    return NEXD_OK;
}

int nexd_mlp_gradients(nexd_mlp *m, const int *idx, int B, float *loss) {
    if (!m || !idx) return m ? mfail(m, NEXD_E_ARG, "null argument") : NEXD_E_ARG;
    MCHK(m, hipSetDevice(m->device));
    int rc = batch_gradients(m, idx, B, loss);
    if (rc) return rc;
    MCHK(m, hipStreamSynchronize(m->stream));
    return NEXD_OK;
}

// nsteps Adam steps over consecutive batches of idx (nsteps * B rows);
// losses[k] (optional) receives the loss of step k before its update.
int nexd_mlp_train(nexd_mlp *m, const int *idx, int B, int nsteps, float lr, float beta1, float beta2, float eps,
                   float *losses) {
    if (!m || !idx || nsteps <= 0) return m ? mfail(m, NEXD_E_ARG, "bad argument") : NEXD_E_ARG;
    MCHK(m, hipSetDevice(m->device));
    MCHK(m, hipEventRecord(m->ev0, m->stream));
    for (int s = 0; s < nsteps; ++s) {
        int rc = batch_gradients(m, idx + (size_t)s * B, B, nullptr);
        if (rc) return rc;
        if (losses) MCHK(m, hipMemcpyAsync(losses + s, m->loss, 4, hipMemcpyDeviceToHost, m->stream));
        adam(m, lr, beta1, beta2, eps);
    }
    MCHK(m, hipEventRecord(m->ev1, m->stream));
    MCHK(m, hipStreamSynchronize(m->stream));
    MCHK(m, hipGetLastError());
    MCHK(m, hipEventElapsedTime(&m->last_ms, m->ev0, m->ev1));
    return NEXD_OK;
}

int nexd_mlp_last_run_ms(nexd_mlp *m, double *ms) {
    if (!m || !ms) return NEXD_E_ARG;
    *ms = m->last_ms;
    return NEXD_OK;
}

void nexd_mlp_destroy(nexd_mlp *m) {
    if (!m) return;
    if (m->stream) (void)hipStreamSynchronize(m->stream);
    for (void *p : m->allocs) (void)hipFree(p);
    if (m->dx) (void)hipFree(m->dx);
    if (m->dt) (void)hipFree(m->dt);
    if (m->ev0) (void)hipEventDestroy(m->ev0);
    if (m->ev1) (void)hipEventDestroy(m->ev1);
    if (m->stream) (void)hipStreamDestroy(m->stream);
    delete m;
}

const char *nexd_mlp_last_error(const nexd_mlp *m) { return m ? m->err.c_str() : g_mlp_create_error.c_str(); }

}  // extern "C"
