// lee_mlp.hip -- the learned ground model: a small MLP (4 -> 64 -> 64 -> 64 -> 8, GELU) that
// maps ground parameters to the residues of the ground impedance's poles, trained and run on
// the GPU with BF16 MFMA products and FP32 accumulation, FP32 master weights and Adam.
//
// Every product of the network is one kernel, C[M x N] = A[M x K] B[K x N] with strided
// operands (forward: activations x weights^T; backward: dZ^T x activations and dZ x weights),
// one 64-lane wave per 16 x 16 tile of C, gfx950's v_mfma_f32_16x16x32_bf16 over K in steps
// of 32 (and v_mfma_f32_16x16x16_bf16 for a last step of 16 when K is an odd multiple of 16);
// the FP32 operands are rounded to BF16 (nearest even, NaN kept) as they are loaded.  Elementwise
// kernels do the bias, GELU and its derivative, the loss gradient, the bias sums and Adam.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../include/lee_abi.h"

namespace {

typedef short v4s __attribute__((ext_vector_type(4)));
typedef short v8s __attribute__((ext_vector_type(8)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef float v4f __attribute__((ext_vector_type(4)));

// FP32 -> BF16 bits, round to nearest even; a NaN stays a (quiet) NaN instead of rounding
// into the exponent field.
__device__ __forceinline__ short to_bf16(float x)
{
    unsigned u = __float_as_uint(x);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (short)((u >> 16) | 0x0040u);
    u = u + 0x7fffu + ((u >> 16) & 1u);
    return (short)(u >> 16);
}

// C(m, n) = sum_k A(m, k) B(k, n) [+ bias(n)]; A(m, k) = pa[m*am + k*ak], B(k, n) = pb[k*bk + n*bn].
// M, N, K multiples of 16.  Block: 4 waves, one 16x16 tile each.  With r = l % 16 and
// q = l / 16, lane l of a wave holds A(r, k0+8q .. k0+8q+7) and B(k0+8q .. +7, r) in a step of
// 32 (A(r, k0+4q .. +3), B(k0+4q .. +3, r) in a step of 16) and receives C(4q .. 4q+3, r).
__global__ __launch_bounds__(256) void k_gemm_bf16(const float *__restrict__ pa, long am, long ak,
                                                   const float *__restrict__ pb, long bk, long bn,
                                                   float *__restrict__ pc, long cm, const float *__restrict__ bias,
                                                   int M, int N, int K)
{
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int tiles_n = N >> 4;
    const int tile = blockIdx.x * 4 + wave;
    if (tile >= (M >> 4) * tiles_n) return;
    const int m0 = (tile / tiles_n) << 4, n0 = (tile % tiles_n) << 4;
    const int r = lane & 15, q = lane >> 4;
    v4f acc = {0.f, 0.f, 0.f, 0.f};
    int k0 = 0;
    for (; k0 + 32 <= K; k0 += 32) {
        v8s a, b;
#pragma unroll
        for (int t = 0; t < 8; ++t) {
            const int k = k0 + 8 * q + t;
            a[t] = to_bf16(pa[(long)(m0 + r) * am + (long)k * ak]);
            b[t] = to_bf16(pb[(long)k * bk + (long)(n0 + r) * bn]);
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(v8bf, a), __builtin_bit_cast(v8bf, b), acc, 0, 0, 0);
    }
    if (k0 < K) {
        v4s a, b;
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const int k = k0 + 4 * q + t;
            a[t] = to_bf16(pa[(long)(m0 + r) * am + (long)k * ak]);
            b[t] = to_bf16(pb[(long)k * bk + (long)(n0 + r) * bn]);
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
    }
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        const int m = m0 + 4 * q + t, n = n0 + r;
        pc[(long)m * cm + n] = acc[t] + (bias ? bias[n] : 0.f);
    }
}

__device__ __forceinline__ float gelu(float x)
{
    const float c = 0.7978845608028654f;  // sqrt(2/pi)
    return 0.5f * x * (1.f + tanhf(c * (x + 0.044715f * x * x * x)));
}
__device__ __forceinline__ float gelu_d(float x)
{
    const float c = 0.7978845608028654f;
    const float u = c * (x + 0.044715f * x * x * x);
    const float th = tanhf(u);
    return 0.5f * (1.f + th) + 0.5f * x * (1.f - th * th) * c * (1.f + 3.f * 0.044715f * x * x);
}

__global__ void k_gelu(const float *z, float *a, long n)
{
    for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) a[t] = gelu(z[t]);
}
// dz = da * gelu'(z)
__global__ void k_gelu_back(const float *z, const float *da, float *dz, long n)
{
    for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x)
        dz[t] = da[t] * gelu_d(z[t]);
}
// loss gradient of the mean squared error over the valid rows (< n) and columns (< nout) of a
// [rows x ld] output: dy = 2 (y - t) / (n nout), zero elsewhere; loss partial sums per block
__global__ void k_mse_back(const float *y, const float *t, float *dy, int rows, int ld, int n, int nout, float *loss)
{
    __shared__ float part[256];
    float s = 0.f;
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < (long)rows * ld; e += (long)gridDim.x * blockDim.x) {
        const int i = (int)(e / ld), j = (int)(e % ld);
        float g = 0.f;
        if (i < n && j < nout) {
            const float d = y[e] - t[(long)i * nout + j];
            s += d * d;
            g = 2.f * d / ((float)n * nout);
        }
        dy[e] = g;
    }
    part[threadIdx.x] = s;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) loss[blockIdx.x] = part[0];
}
// db(n) = sum over rows of dz(row, n): one 64-lane wave per column, lanes stride the rows,
// then a butterfly over the wave
__global__ __launch_bounds__(256) void k_colsum(const float *dz, float *db, int rows, int cols)
{
    const int n = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (n >= cols) return;
    float s = 0.f;
    for (int i = lane; i < rows; i += 64) s += dz[(long)i * cols + n];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if (lane == 0) db[n] = s;
}
// Adam on FP32 parameters (m, v moments), step t (1-based)
__global__ void k_adam(float *w, const float *g, float *m, float *v, long n, float lr, float b1, float b2, float eps,
                       float c1, float c2)
{
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
        const float gg = g[e];
        const float mm = b1 * m[e] + (1.f - b1) * gg;
        const float vv = b2 * v[e] + (1.f - b2) * gg * gg;
        m[e] = mm;
        v[e] = vv;
        w[e] -= lr * (mm / c1) / (sqrtf(vv / c2) + eps);
    }
}

unsigned blocks_for(long n) { long b = (n + 255) / 256; return (unsigned)(b > 4096 ? 4096 : (b ? b : 1)); }

}  // namespace

// ------------------------------------------------------------------------------------
// Network object
// ------------------------------------------------------------------------------------
struct lee_mlp {
    static constexpr int L = 4;
    int din[L] = {16, 64, 64, 64};    // padded inputs of each layer (4 real inputs)
    int dout[L] = {64, 64, 64, 16};   // padded outputs (8 real outputs)
    int n_in = 4, n_out = 8;
    hipStream_t stream = nullptr;
    std::string err;
    // parameters (FP32 master), gradients, Adam moments: one flat buffer each, layer offsets
    std::vector<long> woff, boff;
    long nparam = 0;
    float *w = nullptr, *g = nullptr, *m1 = nullptr, *m2 = nullptr;
    long step = 0;
    // activations for a batch capacity
    int cap = 0;
    float *x = nullptr, *t = nullptr, *dy = nullptr, *loss = nullptr;
    float *z[L] = {}, *a[L] = {}, *dz[L] = {}, *da[L] = {};
    std::vector<void *> allocs;
    ~lee_mlp()
    {
        if (stream) (void)hipStreamSynchronize(stream);
        for (void *p : allocs) (void)hipFree(p);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

int mlp_err(lee_mlp *n, int code, const std::string &m)
{
    if (n) n->err = m;
    return code;
}
#define MHIP(n, expr)                                                                          \
    do {                                                                                       \
        hipError_t e_ = (expr);                                                                \
        if (e_ != hipSuccess) return mlp_err((n), LEE_ERR_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

int mlp_alloc(lee_mlp *n, float **p, long count)
{
    void *q = nullptr;
    MHIP(n, hipMalloc(&q, (size_t)std::max(count, 1L) * sizeof(float)));
    MHIP(n, hipMemsetAsync(q, 0, (size_t)std::max(count, 1L) * sizeof(float), n->stream));
    n->allocs.push_back(q);
    *p = (float *)q;
    return LEE_OK;
}

// Give back a buffer of the batch (the stream must be idle).
void mlp_release(lee_mlp *n, float *&p)
{
    if (!p) return;
    for (auto it = n->allocs.begin(); it != n->allocs.end(); ++it)
        if (*it == (void *)p) {
            n->allocs.erase(it);
            break;
        }
    (void)hipFree(p);
    p = nullptr;
}

// Batch buffers for `rows` rows: kept while they are large enough; on growth the old set is
// freed and the new one sized with headroom (1.5x), so a run of growing batches reallocates
// a few times, not once per call.
int mlp_batch(lee_mlp *n, int rows)
{
    if (rows <= n->cap) return LEE_OK;
    int rc;
    const long want = n->cap > 0 ? std::max((long)rows, (long)n->cap * 3 / 2) : (long)rows;
    const int R = (int)((want + 15) / 16 * 16);
    MHIP(n, hipStreamSynchronize(n->stream));
    mlp_release(n, n->x); mlp_release(n, n->t); mlp_release(n, n->dy); mlp_release(n, n->loss);
    for (int l = 0; l < lee_mlp::L; ++l) {
        mlp_release(n, n->z[l]); mlp_release(n, n->a[l]); mlp_release(n, n->dz[l]); mlp_release(n, n->da[l]);
    }
    n->cap = 0;
    if ((rc = mlp_alloc(n, &n->x, (long)R * n->din[0])) || (rc = mlp_alloc(n, &n->t, (long)R * n->n_out)) ||
        (rc = mlp_alloc(n, &n->dy, (long)R * n->dout[lee_mlp::L - 1])) || (rc = mlp_alloc(n, &n->loss, 4096)))
        return rc;
    for (int l = 0; l < lee_mlp::L; ++l)
        if ((rc = mlp_alloc(n, &n->z[l], (long)R * n->dout[l])) || (rc = mlp_alloc(n, &n->a[l], (long)R * n->dout[l])) ||
            (rc = mlp_alloc(n, &n->dz[l], (long)R * n->dout[l])) || (rc = mlp_alloc(n, &n->da[l], (long)R * n->din[l])))
            return rc;
    n->cap = R;
    return LEE_OK;
}

void gemm(lee_mlp *n, const float *A, long am, long ak, const float *B, long bk, long bn, float *C, long cm,
          const float *bias, int M, int N, int K)
{
    const int tiles = (M / 16) * (N / 16);
    hipLaunchKernelGGL(k_gemm_bf16, dim3((tiles + 3) / 4), dim3(256), 0, n->stream, A, am, ak, B, bk, bn, C, cm, bias, M, N, K);
}

// inputs [rows x n_in] (host) -> padded device batch
int mlp_load_x(lee_mlp *n, const float *x, int rows)
{
    const int R = (rows + 15) / 16 * 16;
    std::vector<float> h((size_t)R * n->din[0], 0.f);
    for (int i = 0; i < rows; ++i)
        for (int j = 0; j < n->n_in; ++j) h[(size_t)i * n->din[0] + j] = x[(size_t)i * n->n_in + j];
    MHIP(n, hipMemcpyAsync(n->x, h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice, n->stream));
    MHIP(n, hipStreamSynchronize(n->stream));
    return LEE_OK;
}

void mlp_forward(lee_mlp *n, int R)
{
    const float *in = n->x;
    for (int l = 0; l < lee_mlp::L; ++l) {
        const float *W = n->w + n->woff[l], *b = n->w + n->boff[l];
        gemm(n, in, n->din[l], 1, W, 1, n->din[l], n->z[l], n->dout[l], b, R, n->dout[l], n->din[l]);
        const long cnt = (long)R * n->dout[l];
        if (l < lee_mlp::L - 1) {
            hipLaunchKernelGGL(k_gelu, dim3(blocks_for(cnt)), dim3(256), 0, n->stream, n->z[l], n->a[l], cnt);
            in = n->a[l];
        }
    }
}

// backward from dy (already in n->dy) into n->g
void mlp_backward(lee_mlp *n, int R)
{
    const float *dcur = n->dy;
    for (int l = lee_mlp::L - 1; l >= 0; --l) {
        const int o = n->dout[l], i = n->din[l];
        const float *dzl = n->dy;  // linear output layer: dZ = dY
        if (l < lee_mlp::L - 1) {
            const long cnt = (long)R * o;
            hipLaunchKernelGGL(k_gelu_back, dim3(blocks_for(cnt)), dim3(256), 0, n->stream, n->z[l], dcur, n->dz[l], cnt);
            dzl = n->dz[l];
        }
        const float *prev = l ? n->a[l - 1] : n->x;
        // dW [o x i] = dZ^T [o x R] * prev [R x i]
        gemm(n, dzl, 1, o, prev, i, 1, n->g + n->woff[l], i, nullptr, o, i, R);
        hipLaunchKernelGGL(k_colsum, dim3((o + 3) / 4), dim3(256), 0, n->stream, dzl, n->g + n->boff[l], R, o);
        if (l > 0) {
            // dA_prev [R x i] = dZ [R x o] * W [o x i]
            gemm(n, dzl, o, 1, n->w + n->woff[l], i, 1, n->da[l], i, nullptr, R, i, o);
            dcur = n->da[l];
        }
    }
}

// loss gradient for targets t [rows x n_out] (host); returns the mean squared error
int mlp_loss(lee_mlp *n, const float *t, int rows, double *mse)
{
    const int R = (rows + 15) / 16 * 16;
    MHIP(n, hipMemcpyAsync(n->t, t, (size_t)rows * n->n_out * sizeof(float), hipMemcpyHostToDevice, n->stream));
    const int last = lee_mlp::L - 1;
    const unsigned nb = 64;
    hipLaunchKernelGGL(k_mse_back, dim3(nb), dim3(256), 0, n->stream, n->z[last], n->t, n->dy, R, n->dout[last], rows,
                       n->n_out, n->loss);
    std::vector<float> part(nb);
    MHIP(n, hipMemcpyAsync(part.data(), n->loss, nb * sizeof(float), hipMemcpyDeviceToHost, n->stream));
    MHIP(n, hipStreamSynchronize(n->stream));
    double s = 0;
    for (float v : part) s += v;
    if (mse) *mse = s / ((double)rows * n->n_out);
    return LEE_OK;
}

}  // namespace

extern "C" {

int lee_mlp_create(int device, unsigned long long seed, lee_mlp **out)
{
    if (!out) return LEE_ERR_ARG;
    *out = nullptr;
    std::unique_ptr<lee_mlp> n(new lee_mlp());
    lee_mlp *N = n.get();
    MHIP(N, hipSetDevice(device));
    MHIP(N, hipStreamCreateWithFlags(&N->stream, hipStreamNonBlocking));
    long off = 0;
    for (int l = 0; l < lee_mlp::L; ++l) {
        N->woff.push_back(off); off += (long)N->dout[l] * N->din[l];
        N->boff.push_back(off); off += N->dout[l];
    }
    N->nparam = off;
    int rc;
    if ((rc = mlp_alloc(N, &N->w, off)) || (rc = mlp_alloc(N, &N->g, off)) || (rc = mlp_alloc(N, &N->m1, off)) ||
        (rc = mlp_alloc(N, &N->m2, off)))
        return rc;
    // He-uniform initial weights over the real fan-in; padded rows and columns stay zero
    std::vector<float> h(off, 0.f);
    unsigned long long st = seed * 6364136223846793005ULL + 1442695040888963407ULL;
    auto uni = [&]() {
        st = st * 6364136223846793005ULL + 1442695040888963407ULL;
        return (double)(st >> 11) * (1.0 / 9007199254740992.0);
    };
    const int real_in[lee_mlp::L] = {N->n_in, 64, 64, 64}, real_out[lee_mlp::L] = {64, 64, 64, N->n_out};
    for (int l = 0; l < lee_mlp::L; ++l) {
        const double lim = std::sqrt(6.0 / real_in[l]) * (l == lee_mlp::L - 1 ? 0.1 : 1.0);
        for (int o = 0; o < real_out[l]; ++o)
            for (int i = 0; i < real_in[l]; ++i) h[N->woff[l] + (long)o * N->din[l] + i] = (float)((2 * uni() - 1) * lim);
    }
    MHIP(N, hipMemcpyAsync(N->w, h.data(), off * sizeof(float), hipMemcpyHostToDevice, N->stream));
    MHIP(N, hipStreamSynchronize(N->stream));
    *out = n.release();
    return LEE_OK;
}

int lee_mlp_destroy(lee_mlp *n)
{
    delete n;
    return LEE_OK;
}

const char *lee_mlp_error(const lee_mlp *n) { return n ? n->err.c_str() : ""; }

long long lee_mlp_param_count(const lee_mlp *n) { return n ? n->nparam : 0; }

int lee_mlp_get_params(lee_mlp *n, float *w, long long cap)
{
    if (!n || !w || cap < n->nparam) return mlp_err(n, LEE_ERR_ARG, "buffer too small");
    MHIP(n, hipMemcpyAsync(w, n->w, n->nparam * sizeof(float), hipMemcpyDeviceToHost, n->stream));
    MHIP(n, hipStreamSynchronize(n->stream));
    return LEE_OK;
}

int lee_mlp_set_params(lee_mlp *n, const float *w, long long count)
{
    if (!n || !w || count != n->nparam) return mlp_err(n, LEE_ERR_ARG, "parameter count mismatch");
    MHIP(n, hipMemcpyAsync(n->w, w, n->nparam * sizeof(float), hipMemcpyHostToDevice, n->stream));
    MHIP(n, hipStreamSynchronize(n->stream));
    return LEE_OK;
}

int lee_mlp_predict(lee_mlp *n, const float *x, int rows, float *y)
{
    if (!n || !x || !y || rows <= 0) return mlp_err(n, LEE_ERR_ARG, "bad arguments");
    int rc;
    if ((rc = mlp_batch(n, rows)) || (rc = mlp_load_x(n, x, rows))) return rc;
    const int R = (rows + 15) / 16 * 16, last = lee_mlp::L - 1;
    mlp_forward(n, R);
    std::vector<float> h((size_t)R * n->dout[last]);
    MHIP(n, hipMemcpyAsync(h.data(), n->z[last], h.size() * sizeof(float), hipMemcpyDeviceToHost, n->stream));
    MHIP(n, hipStreamSynchronize(n->stream));
    for (int i = 0; i < rows; ++i)
        for (int j = 0; j < n->n_out; ++j) y[(size_t)i * n->n_out + j] = h[(size_t)i * n->dout[last] + j];
    return LEE_OK;
}

int lee_mlp_gradient(lee_mlp *n, const float *x, const float *t, int rows, float *grad, double *mse)
{
    if (!n || !x || !t || rows <= 0) return mlp_err(n, LEE_ERR_ARG, "bad arguments");
    int rc;
    if ((rc = mlp_batch(n, rows)) || (rc = mlp_load_x(n, x, rows))) return rc;
    const int R = (rows + 15) / 16 * 16;
    mlp_forward(n, R);
    if ((rc = mlp_loss(n, t, rows, mse))) return rc;
    mlp_backward(n, R);
    MHIP(n, hipGetLastError());
    if (grad) {
        MHIP(n, hipMemcpyAsync(grad, n->g, n->nparam * sizeof(float), hipMemcpyDeviceToHost, n->stream));
        MHIP(n, hipStreamSynchronize(n->stream));
    }
    return LEE_OK;
}

int lee_mlp_train(lee_mlp *n, const float *x, const float *t, int rows, int batch, int epochs, float lr,
                  unsigned long long seed, double *final_mse)
{
    if (!n || !x || !t || rows <= 0 || batch <= 0 || epochs < 0) return mlp_err(n, LEE_ERR_ARG, "bad arguments");
    batch = std::min(batch, rows);
    int rc;
    if ((rc = mlp_batch(n, batch))) return rc;
    std::vector<int> order(rows);
    for (int i = 0; i < rows; ++i) order[i] = i;
    unsigned long long st = seed + 0x9E3779B97F4A7C15ULL;
    std::vector<float> bx((size_t)batch * n->n_in), bt((size_t)batch * n->n_out);
    double mse = 0;
    for (int ep = 0; ep < epochs; ++ep) {
        for (int i = rows - 1; i > 0; --i) {  // Fisher-Yates with a 64-bit LCG
            st = st * 6364136223846793005ULL + 1442695040888963407ULL;
            std::swap(order[i], order[(int)((st >> 33) % (unsigned long long)(i + 1))]);
        }
        for (int b0 = 0; b0 + batch <= rows; b0 += batch) {
            for (int r = 0; r < batch; ++r) {
                std::memcpy(&bx[(size_t)r * n->n_in], &x[(size_t)order[b0 + r] * n->n_in], n->n_in * sizeof(float));
                std::memcpy(&bt[(size_t)r * n->n_out], &t[(size_t)order[b0 + r] * n->n_out], n->n_out * sizeof(float));
            }
            if ((rc = lee_mlp_gradient(n, bx.data(), bt.data(), batch, nullptr, &mse))) return rc;
            n->step += 1;
            const float b1 = 0.9f, b2 = 0.999f;
            const float c1 = 1.f - std::pow(b1, (float)n->step), c2 = 1.f - std::pow(b2, (float)n->step);
            hipLaunchKernelGGL(k_adam, dim3(blocks_for(n->nparam)), dim3(256), 0, n->stream, n->w, n->g, n->m1, n->m2,
                               (long)n->nparam, lr, b1, b2, 1e-8f, c1, c2);
        }
    }
    MHIP(n, hipStreamSynchronize(n->stream));
    if (final_mse) *final_mse = mse;
    return LEE_OK;
}

}  // extern "C"
