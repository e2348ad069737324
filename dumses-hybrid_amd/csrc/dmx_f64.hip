// float64 engine: the reference's arithmetic, compiled with -ffp-contract=off
// so every kernel reproduces the reference bit for bit.
#define DMX_T double
#define DMX_NS f64
#include "dmx_step.inc"

#include <mutex>

namespace dmx {
Engine* make_engine_f64(const dmx_params& p, int* rc)
{
    f64::EngineT* e = new f64::EngineT();
    *rc = e->init(p);
    if (*rc) {
        delete e;
        return nullptr;
    }
    return e;
}
Batch* make_batch_f64(Engine* const* es, int n, int nstreams, int* rc)
{
    f64::BatchT* b = new f64::BatchT();
    *rc = b->init(es, n, nstreams);
    if (*rc) {
        delete b;
        return nullptr;
    }
    return b;
}
} // namespace dmx

// ---------------------------------------------------------------------------
// dmx_compute_dt_noalloc: the CFL timestep (physics.f90:18-132) of a float64
// state that the caller already holds in device memory in the reference's
// uin layout, enqueued on the caller's stream with no device allocation.
// The block minima of the first pass go to `scratch`, a file-scope device
// array of kNoAllocPartials doubles (16 KiB, held by every device the
// library runs on until the process ends).  Every call on a device shares
// that device's array and keeps the last call's minima in it: calls that
// overlap on one device race.  The first pass writes every partial the
// second pass reads, so nothing needs clearing between calls.
// ---------------------------------------------------------------------------
namespace dmx {
namespace f64 {

constexpr int kNoAllocPartials = 2048;
static __device__ __align__(256) double scratch[kNoAllocPartials];

// the device address of `scratch` on each device, looked up once per
// device with that device current (the array's name is a host-side
// stand-in, not a device address)
constexpr int kMaxDevices = 64;
static double* g_scratch_ptr[kMaxDevices];
static std::once_flag g_scratch_once[kMaxDevices];
static hipError_t g_scratch_err[kMaxDevices];

static hipError_t scratch_ptr(int dev, double** out)
{
    std::call_once(g_scratch_once[dev], [dev] {
        void* p = nullptr;
        g_scratch_err[dev] = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        g_scratch_ptr[dev] = (double*)p;
    });
    *out = g_scratch_ptr[dev];
    return g_scratch_err[dev];
}

// pass 1: grid-stride over the interior cells, one minimum per block into
// part[blockIdx.x]; the cell field of each cell is the mean of its left and
// right face fields uin(:,:,:,6..8) and (:,:,:,9..11), as primitive forms it
// (scheme.f90:55-57)
__global__ __launch_bounds__(256) void k_dt_uin_min(Step st, const double* __restrict__ uin, double* part)
{
    const Geom& g = st.g;
    const long long gx = g.nx + 6, gyy = g.ny + 6;
    const long long vs = gx * gyy * (long long)(g.nz + 6);
    const long long ncell = (long long)g.nx * g.ny * g.nz;
    double best = 1e300;
    for (long long n = (long long)blockIdx.x * 256 + threadIdx.x; n < ncell; n += (long long)gridDim.x * 256) {
        const int i = (int)(n % g.nx) + 1;
        const int j = (int)((n / g.nx) % g.ny) + 1;
        const int k = (int)(n / ((long long)g.nx * g.ny)) + 1;
        const double* u = uin + ((i + 2) + gx * ((j + 2) + gyy * (long long)(k + 2)));
        best = fmn(best, dt_cell(st, u[0], u[vs], u[2 * vs], u[3 * vs], u[4 * vs], u[5 * vs], u[8 * vs], u[6 * vs],
                                 u[9 * vs], u[7 * vs], u[10 * vs]));
    }
    block_min_out(best, part + blockIdx.x);
}

// pass 2: dt = courant * min(courant*dx/smallc, block minima)
__global__ __launch_bounds__(256) void k_dt_uin_final(const double* part, int nb, double init, double courant,
                                                      double* dt)
{
    dt_final_body(dt, part, nb, init, courant, nullptr);
}

int compute_dt_noalloc(const dmx_params& p, const double* uin, double* dt, hipStream_t stream)
{
    int dev = -1;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) {
        set_error("dmx_compute_dt_noalloc: device %d beyond the %d devices it supports", dev, kMaxDevices);
        return DMX_E_PARAM;
    }
    double* part = nullptr;
    HIPCHK(scratch_ptr(dev, &part));
    Step st;
    memset(&st, 0, sizeof(st));
    Geom& g = st.g;
    g.nx = p.nx; g.ny = p.ny; g.nz = p.nz;
    g.dx = (p.xmax - p.xmin) / p.nx;
    g.dy = (p.ymax - p.ymin) / p.ny;
    g.dz = (p.zmax - p.zmin) / p.nz;
    g.xmin = p.xmin; g.xmax = p.xmax; g.ymin = p.ymin; g.ymax = p.ymax; g.zmin = p.zmin; g.zmax = p.zmax;
    st.ph.gamma = p.gamma; st.ph.ciso = p.ciso; st.ph.omega0 = p.omega0; st.ph.courant = p.courant;
    st.ph.iso = p.iso; st.ph.fargo = p.fargo;
    const long long ncell = (long long)p.nx * p.ny * p.nz;
    const long long want = (ncell + 255) / 256;
    const int nb = (int)(want < kNoAllocPartials ? want : kNoAllocPartials);
    hipLaunchKernelGGL(k_dt_uin_min, dim3(nb), dim3(256), 0, stream, st, uin, part);
    hipLaunchKernelGGL(k_dt_uin_final, dim3(1), dim3(256), 0, stream, (const double*)part, nb,
                       p.courant * g.dx / 1e-10, p.courant, dt);
    HIPCHK(hipGetLastError());
    return DMX_OK;
}

} // namespace f64

int noalloc_compute_dt_f64(const dmx_params& p, const double* uin, double* dt, void* stream)
{
    return f64::compute_dt_noalloc(p, uin, dt, (hipStream_t)stream);
}
} // namespace dmx

