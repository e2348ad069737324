// dmx_mlp.hip -- training of the learned slope limiter on the GPU:
// forward and backward passes of the 40-64-64-8 GELU network (float32),
// mean-squared-error loss, Adam, and on-device minibatch sampling from a
// resident data set; plus a self-test of the bf16 MFMA lane layout that the
// inference kernel of the step relies on.
#include <hip/hip_runtime.h>

#include <stdint.h>
#include <string.h>

#include <vector>

#include "dmx.h"
#include "dmx_engine.h"
#include "dmx_mlp.h"

namespace dmx {
namespace mlp {

// pre[s][o] = b[o] + sum_i in[s][i] W[i][o]; act = gelu(pre) when act != 0
__global__ __launch_bounds__(256) void k_dense_fwd(const float* in, int nin, const float* W, const float* b,
                                                   int nout, int n, float* pre, float* act)
{
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)n * nout) return;
    int s = (int)(t / nout), o = (int)(t % nout);
    const float* x = in + (long long)s * nin;
    float a = b[o];
    for (int i = 0; i < nin; i++) a += x[i] * W[i * nout + o];
    pre[t] = a;
    if (act) act[t] = gelu(a);
}

// g = (y - T) / n and the per-block partial sums of 0.5 (y - T)^2 / n
__global__ __launch_bounds__(256) void k_out_grad(const float* y, const float* T, int n, float* g, float* part)
{
    __shared__ float red[256];
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    float l = 0.0f;
    if (t < (long long)n * NOUT) {
        float r = y[t] - T[t];
        g[t] = r / (float)n;
        l = 0.5f * r * r / (float)n;
    }
    red[threadIdx.x] = l;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// gW[i][o] = sum_s in[s][i] g[s][o] (row i = nin gives the bias gradient):
// one block per output column o, threads over (i, sample chunk)
__global__ __launch_bounds__(256) void k_dense_wgrad(const float* in, int nin, const float* g, int nout, int n,
                                                     float* gW, float* gb)
{
    __shared__ float red[256];
    const int o = blockIdx.x;
    for (int i = 0; i <= nin; i++) {
        float a = 0.0f;
        for (int s = threadIdx.x; s < n; s += 256) {
            float x = i < nin ? in[(long long)s * nin + i] : 1.0f;
            a += x * g[(long long)s * nout + o];
        }
        red[threadIdx.x] = a;
        __syncthreads();
        for (int w = 128; w > 0; w >>= 1) {
            if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
            __syncthreads();
        }
        if (threadIdx.x == 0) {
            if (i < nin) gW[i * nout + o] = red[0];
            else gb[o] = red[0];
        }
        __syncthreads();
    }
}

// gin[s][i] = gelu'(pre_in[s][i]) sum_o g[s][o] W[i][o]
__global__ __launch_bounds__(256) void k_dense_igrad(const float* g, int nout, const float* W, int nin, int n,
                                                     const float* pre_in, float* gin)
{
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)n * nin) return;
    int s = (int)(t / nin), i = (int)(t % nin);
    const float* gs = g + (long long)s * nout;
    float a = 0.0f;
    for (int o = 0; o < nout; o++) a += gs[o] * W[i * nout + o];
    gin[t] = a * gelu_grad(pre_in[t]);
}

// Adam (Kingma and Ba 2015), bias-corrected step size
__global__ __launch_bounds__(256) void k_adam(float* th, const float* gr, float* m, float* v, int n, float lr,
                                              float b1, float b2, float eps, float c1, float c2)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    float g = gr[t];
    float mm = b1 * m[t] + (1.0f - b1) * g;
    float vv = b2 * v[t] + (1.0f - b2) * g * g;
    m[t] = mm;
    v[t] = vv;
    th[t] -= lr * (mm / c1) / (sqrtf(vv / c2) + eps);
}

// minibatch of B rows drawn from n rows with a counter-based hash
// (splitmix64 finaliser of seed + draw number)
__device__ __forceinline__ unsigned long long mix64(unsigned long long z)
{
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

__global__ __launch_bounds__(256) void k_gather(const float* X, const float* T, long long n, int B,
                                                unsigned long long key, float* Xb, float* Tb)
{
    int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B) return;
    long long r = (long long)(mix64(key + (unsigned long long)b * 0x9e3779b97f4a7c15ULL) % (unsigned long long)n);
    for (int i = 0; i < NIN; i++) Xb[(long long)b * NIN + i] = X[r * NIN + i];
    for (int o = 0; o < NOUT; o++) Tb[(long long)b * NOUT + o] = T[r * NOUT + o];
}

// D = A B for A[i][k] = (i + 2k mod 7) / 8, B[k][j] = (3k + j mod 5) / 4 - 0.5
// (exact in bf16), through the MFMA lane layout of dmx_mlp.h
__global__ void k_mfma_selftest(float* out)
{
    const int l = threadIdx.x;
    bf16x8 a, b;
    for (int e = 0; e < 8; e++) {
        int k = 8 * (l / 16) + e;
        a[e] = (__bf16)(float((l % 16 + 2 * k) % 7) / 8.0f);
        b[e] = (__bf16)(float((3 * k + l % 16) % 5) / 4.0f - 0.5f);
    }
    f32x4 c = {0.0f, 0.0f, 0.0f, 0.0f};
    c = mfma16x16x32(a, b, c);
    for (int r = 0; r < 4; r++) out[(4 * (l / 16) + r) * 16 + (l % 16)] = c[r];
}

} // namespace mlp
} // namespace dmx

using namespace dmx::mlp;

struct dmx_mlp {
    int cap = 0;
    float* th = nullptr;     // parameters
    float* gr = nullptr;     // gradient
    float* m = nullptr;
    float* v = nullptr;
    float* ws = nullptr;     // per-batch workspace
    float* X = nullptr;      // resident data set
    float* T = nullptr;
    long long ndata = 0;
    long long adam_t = 0;
    hipStream_t stream = nullptr;
    // workspace views
    float *xb, *tb, *a1, *h1, *a2, *h2, *y, *g3, *g2, *g1, *part;
};

#define MCHK(x)                                                                           \
    do {                                                                                  \
        hipError_t e_ = (x);                                                              \
        if (e_ != hipSuccess) {                                                           \
            dmx::set_error("%s:%d %s: %s", __FILE__, __LINE__, #x, hipGetErrorString(e_)); \
            return DMX_E_HIP;                                                             \
        }                                                                                 \
    } while (0)

static int cdivi(long long a, int b) { return (int)((a + b - 1) / b); }

// forward of n <= cap rows already in m->xb; leaves a1, h1, a2, h2, y
static int fwd(dmx_mlp* m, int n)
{
    const float* W1 = m->th + OFF_W1; const float* b1 = m->th + OFF_B1;
    const float* W2 = m->th + OFF_W2; const float* b2 = m->th + OFF_B2;
    const float* W3 = m->th + OFF_W3; const float* b3 = m->th + OFF_B3;
    hipLaunchKernelGGL(k_dense_fwd, dim3(cdivi((long long)n * NH, 256)), dim3(256), 0, m->stream, m->xb, NIN, W1, b1,
                       NH, n, m->a1, m->h1);
    hipLaunchKernelGGL(k_dense_fwd, dim3(cdivi((long long)n * NH, 256)), dim3(256), 0, m->stream, m->h1, NH, W2, b2,
                       NH, n, m->a2, m->h2);
    hipLaunchKernelGGL(k_dense_fwd, dim3(cdivi((long long)n * NOUT, 256)), dim3(256), 0, m->stream, m->h2, NH, W3,
                       b3, NOUT, n, m->y, (float*)nullptr);
    MCHK(hipGetLastError());
    return DMX_OK;
}

// loss and gradient of the rows in xb/tb (after fwd); loss stays on device
static int bwd(dmx_mlp* m, int n)
{
    const float* W2 = m->th + OFF_W2;
    const float* W3 = m->th + OFF_W3;
    int nb = cdivi((long long)n * NOUT, 256);
    hipLaunchKernelGGL(k_out_grad, dim3(nb), dim3(256), 0, m->stream, m->y, m->tb, n, m->g3, m->part);
    hipLaunchKernelGGL(k_dense_wgrad, dim3(NOUT), dim3(256), 0, m->stream, m->h2, NH, m->g3, NOUT, n,
                       m->gr + OFF_W3, m->gr + OFF_B3);
    hipLaunchKernelGGL(k_dense_igrad, dim3(cdivi((long long)n * NH, 256)), dim3(256), 0, m->stream, m->g3, NOUT, W3,
                       NH, n, m->a2, m->g2);
    hipLaunchKernelGGL(k_dense_wgrad, dim3(NH), dim3(256), 0, m->stream, m->h1, NH, m->g2, NH, n, m->gr + OFF_W2,
                       m->gr + OFF_B2);
    hipLaunchKernelGGL(k_dense_igrad, dim3(cdivi((long long)n * NH, 256)), dim3(256), 0, m->stream, m->g2, NH, W2,
                       NH, n, m->a1, m->g1);
    hipLaunchKernelGGL(k_dense_wgrad, dim3(NH), dim3(256), 0, m->stream, m->xb, NIN, m->g1, NH, n, m->gr + OFF_W1,
                       m->gr + OFF_B1);
    MCHK(hipGetLastError());
    return DMX_OK;
}

static float sum_parts(dmx_mlp* m, int nparts, int* rc)
{
    std::vector<float> h(nparts);
    if (hipMemcpyAsync(h.data(), m->part, sizeof(float) * nparts, hipMemcpyDeviceToHost, m->stream) != hipSuccess ||
        hipStreamSynchronize(m->stream) != hipSuccess) {
        dmx::set_error("loss readback failed");
        *rc = DMX_E_HIP;
        return 0.0f;
    }
    double s = 0.0;
    for (float x : h) s += x;
    *rc = DMX_OK;
    return (float)s;
}

extern "C" {

int dmx_mlp_create(int32_t max_batch, dmx_mlp** out)
{
    if (!out || max_batch < 1 || max_batch > (1 << 22)) {
        dmx::set_error("dmx_mlp_create: max_batch must be in [1, 4194304]");
        return DMX_E_PARAM;
    }
    *out = nullptr;
    dmx_mlp* m = new dmx_mlp();
    m->cap = max_batch;
    size_t cap = (size_t)max_batch;
    size_t wsz = cap * (NIN + NOUT + 4 * NH + NOUT + NOUT + 2 * NH) + cdivi((long long)cap * NOUT, 256) + 64;
    if (hipStreamCreate(&m->stream) != hipSuccess || hipMalloc(&m->th, sizeof(float) * NPARAM) != hipSuccess ||
        hipMalloc(&m->gr, sizeof(float) * NPARAM) != hipSuccess || hipMalloc(&m->m, sizeof(float) * NPARAM) != hipSuccess ||
        hipMalloc(&m->v, sizeof(float) * NPARAM) != hipSuccess || hipMalloc(&m->ws, sizeof(float) * wsz) != hipSuccess) {
        dmx::set_error("dmx_mlp_create: device allocation failed");
        dmx_mlp_destroy(m);
        return DMX_E_NOMEM;
    }
    float* p = m->ws;
    m->xb = p; p += cap * NIN;
    m->tb = p; p += cap * NOUT;
    m->a1 = p; p += cap * NH;
    m->h1 = p; p += cap * NH;
    m->a2 = p; p += cap * NH;
    m->h2 = p; p += cap * NH;
    m->y = p; p += cap * NOUT;
    m->g3 = p; p += cap * NOUT;
    m->g2 = p; p += cap * NH;
    m->g1 = p; p += cap * NH;
    m->part = p;
    hipMemsetAsync(m->th, 0, sizeof(float) * NPARAM, m->stream);
    hipMemsetAsync(m->m, 0, sizeof(float) * NPARAM, m->stream);
    hipMemsetAsync(m->v, 0, sizeof(float) * NPARAM, m->stream);
    if (hipStreamSynchronize(m->stream) != hipSuccess) {
        dmx::set_error("dmx_mlp_create: initialisation failed");
        dmx_mlp_destroy(m);
        return DMX_E_HIP;
    }
    *out = m;
    return DMX_OK;
}

void dmx_mlp_destroy(dmx_mlp* m)
{
    if (!m) return;
    if (m->stream) hipStreamSynchronize(m->stream);
    float* bufs[] = {m->th, m->gr, m->m, m->v, m->ws, m->X, m->T};
    for (float* b : bufs)
        if (b) hipFree(b);
    if (m->stream) hipStreamDestroy(m->stream);
    delete m;
}

int dmx_mlp_set_params(dmx_mlp* m, const float* theta, int32_t n)
{
    if (!m || !theta || n != NPARAM) { dmx::set_error("dmx_mlp_set_params: need %d parameters", NPARAM); return DMX_E_PARAM; }
    MCHK(hipMemcpy(m->th, theta, sizeof(float) * NPARAM, hipMemcpyHostToDevice));
    MCHK(hipMemset(m->m, 0, sizeof(float) * NPARAM));
    MCHK(hipMemset(m->v, 0, sizeof(float) * NPARAM));
    m->adam_t = 0;
    return DMX_OK;
}

int dmx_mlp_get_params(dmx_mlp* m, float* theta, int32_t n)
{
    if (!m || !theta || n != NPARAM) { dmx::set_error("dmx_mlp_get_params: need %d parameters", NPARAM); return DMX_E_PARAM; }
    MCHK(hipStreamSynchronize(m->stream));
    MCHK(hipMemcpy(theta, m->th, sizeof(float) * NPARAM, hipMemcpyDeviceToHost));
    return DMX_OK;
}

int dmx_mlp_forward(dmx_mlp* m, const float* X, int32_t n, float* Y)
{
    if (!m || !X || !Y || n < 1 || n > m->cap) { dmx::set_error("dmx_mlp_forward: bad arguments"); return DMX_E_PARAM; }
    MCHK(hipMemcpyAsync(m->xb, X, sizeof(float) * (size_t)n * NIN, hipMemcpyHostToDevice, m->stream));
    int rc = fwd(m, n);
    if (rc) return rc;
    MCHK(hipMemcpyAsync(Y, m->y, sizeof(float) * (size_t)n * NOUT, hipMemcpyDeviceToHost, m->stream));
    MCHK(hipStreamSynchronize(m->stream));
    return DMX_OK;
}

int dmx_mlp_loss_grad(dmx_mlp* m, const float* X, const float* T, int32_t n, float* loss, float* grad)
{
    if (!m || !X || !T || !loss || !grad || n < 1 || n > m->cap) {
        dmx::set_error("dmx_mlp_loss_grad: bad arguments");
        return DMX_E_PARAM;
    }
    MCHK(hipMemcpyAsync(m->xb, X, sizeof(float) * (size_t)n * NIN, hipMemcpyHostToDevice, m->stream));
    MCHK(hipMemcpyAsync(m->tb, T, sizeof(float) * (size_t)n * NOUT, hipMemcpyHostToDevice, m->stream));
    int rc = fwd(m, n);
    if (!rc) rc = bwd(m, n);
    if (rc) return rc;
    *loss = sum_parts(m, cdivi((long long)n * NOUT, 256), &rc);
    if (rc) return rc;
    MCHK(hipMemcpy(grad, m->gr, sizeof(float) * NPARAM, hipMemcpyDeviceToHost));
    return DMX_OK;
}

int dmx_mlp_load_data(dmx_mlp* m, const float* X, const float* T, int64_t n)
{
    if (!m || !X || !T || n < 1) { dmx::set_error("dmx_mlp_load_data: bad arguments"); return DMX_E_PARAM; }
    if (m->X) hipFree(m->X);
    if (m->T) hipFree(m->T);
    m->X = m->T = nullptr;
    m->ndata = 0;
    if (hipMalloc(&m->X, sizeof(float) * (size_t)n * NIN) != hipSuccess ||
        hipMalloc(&m->T, sizeof(float) * (size_t)n * NOUT) != hipSuccess) {
        dmx::set_error("dmx_mlp_load_data: device allocation failed");
        return DMX_E_NOMEM;
    }
    MCHK(hipMemcpy(m->X, X, sizeof(float) * (size_t)n * NIN, hipMemcpyHostToDevice));
    MCHK(hipMemcpy(m->T, T, sizeof(float) * (size_t)n * NOUT, hipMemcpyHostToDevice));
    m->ndata = n;
    return DMX_OK;
}

int dmx_mlp_train(dmx_mlp* m, int32_t steps, int32_t batch, float lr, uint64_t seed, float* losses)
{
    if (!m || steps < 0 || batch < 1 || batch > m->cap || !(lr > 0.0f)) {
        dmx::set_error("dmx_mlp_train: bad arguments");
        return DMX_E_PARAM;
    }
    if (!m->ndata) { dmx::set_error("dmx_mlp_train: no data loaded"); return DMX_E_STATE; }
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const int nparts = cdivi((long long)batch * NOUT, 256);
    for (int s = 0; s < steps; s++) {
        unsigned long long key = seed * 0x632be59bd9b4e019ULL + (unsigned long long)(m->adam_t + 1) * 0x9e3779b97f4a7c15ULL;
        hipLaunchKernelGGL(k_gather, dim3(cdivi(batch, 256)), dim3(256), 0, m->stream, (const float*)m->X,
                           (const float*)m->T, m->ndata, batch, key, m->xb, m->tb);
        int rc = fwd(m, batch);
        if (!rc) rc = bwd(m, batch);
        if (rc) return rc;
        m->adam_t++;
        float c1 = 1.0f - powf(b1, (float)m->adam_t), c2 = 1.0f - powf(b2, (float)m->adam_t);
        hipLaunchKernelGGL(k_adam, dim3(cdivi(NPARAM, 256)), dim3(256), 0, m->stream, m->th, (const float*)m->gr, m->m,
                           m->v, NPARAM, lr, b1, b2, eps, c1, c2);
        MCHK(hipGetLastError());
        if (losses) {
            losses[s] = sum_parts(m, nparts, &rc);
            if (rc) return rc;
        }
    }
    MCHK(hipStreamSynchronize(m->stream));
    return DMX_OK;
}

int dmx_mfma_selftest(float* out)
{
    if (!out) { dmx::set_error("null out"); return DMX_E_PARAM; }
    float* d = nullptr;
    MCHK(hipMalloc(&d, sizeof(float) * 256));
    hipLaunchKernelGGL(k_mfma_selftest, dim3(1), dim3(64), 0, 0, d);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpy(out, d, sizeof(float) * 256, hipMemcpyDeviceToHost);
    hipFree(d);
    if (e != hipSuccess) { dmx::set_error("mfma selftest: %s", hipGetErrorString(e)); return DMX_E_HIP; }
    return DMX_OK;
}

} // extern "C"
