// dmx_abi.cpp -- the extern "C" entry points of include/dmx.h.
#include <hip/hip_runtime.h>

#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <vector>

#include "dmx_engine.h"

namespace dmx {
static thread_local char g_err[512] = "";
void set_error(const char* fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}
} // namespace dmx

struct dmx_ctx {
    dmx::Engine* e;
    dmx_params p;
};

static int validate(const dmx_params* p)
{
    using dmx::set_error;
    if (!p) { set_error("null params"); return DMX_E_PARAM; }
    if (p->nx < 3 || p->ny < 3 || p->nz < 3) {
        set_error("nx, ny, nz must be >= 3 (periodic ghost copies read cells n-2..n)");
        return DMX_E_PARAM;
    }
    if (p->riemann != DMX_LLF && p->riemann != DMX_HLL && p->riemann != DMX_HLLD && p->riemann != DMX_UPWIND) {
        set_error("unknown 1D Riemann solver id %d", p->riemann);
        return DMX_E_PARAM;
    }
    if (p->riemann2d != DMX_LLF && p->riemann2d != DMX_HLLD && p->riemann2d != DMX_UPWIND && p->riemann2d != DMX_HLLA) {
        set_error("unknown 2D Riemann solver id %d", p->riemann2d);
        return DMX_E_PARAM;
    }
    if (p->slope_type < 0 || p->slope_type > 4) {
        set_error("unknown slope type %d", p->slope_type);
        return DMX_E_PARAM;
    }
    for (int d = 0; d < 3; d++) {
        int b = p->bdtype[d];
        if (b != DMX_BC_PERIODIC && b != DMX_BC_ZERO_GRADIENT && !(b == DMX_BC_SHEARINGBOX && d == 0)) {
            set_error("unsupported boundary type %d in direction %d (analytic boundaries are host callbacks)", b, d);
            return DMX_E_PARAM;
        }
    }
    if (p->bdtype[0] == DMX_BC_SHEARINGBOX && p->bdtype[1] != DMX_BC_PERIODIC) {
        set_error("a shearing-box x boundary needs a periodic y boundary");
        return DMX_E_PARAM;
    }
    if (p->omega0 < 0.0) { set_error("omega0 must be >= 0"); return DMX_E_PARAM; }
    if (p->fargo && !(p->omega0 > 0.0)) { set_error("fargo needs omega0 > 0"); return DMX_E_PARAM; }
    if (!(p->gamma > 1.0)) { set_error("gamma must be > 1"); return DMX_E_PARAM; }
    if (!(p->courant > 0.0)) { set_error("courant must be > 0"); return DMX_E_PARAM; }
    if (!(p->nu >= 0.0) || !(p->eta >= 0.0)) { set_error("nu and eta must be >= 0"); return DMX_E_PARAM; }
    if (!(p->xmax > p->xmin) || !(p->ymax > p->ymin) || !(p->zmax > p->zmin)) {
        set_error("empty domain");
        return DMX_E_PARAM;
    }
    if (p->precision != DMX_FP64 && p->precision != DMX_FP32) {
        set_error("unknown precision %d", p->precision);
        return DMX_E_PARAM;
    }
    {
        // the kernels address one plane of float64 with 32-bit byte offsets
        unsigned long long px = ((unsigned long long)p->nx + 6 + 15) / 16 * 16;
        unsigned long long plane = px * (unsigned long long)(p->ny + 6) * (unsigned long long)(p->nz + 6) * 8ull;
        if (plane >= (1ull << 32)) {
            set_error("domain too large: one float64 plane of (nx+6)(ny+6)(nz+6) cells must stay below 4 GiB "
                      "(about 800^3 cells)");
            return DMX_E_PARAM;
        }
    }
    return DMX_OK;
}

extern "C" {

int dmx_create(const dmx_params* p, dmx_ctx** ctx)
{
    if (!ctx) { dmx::set_error("null ctx"); return DMX_E_PARAM; }
    *ctx = nullptr;
    int rc = validate(p);
    if (rc) return rc;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
        dmx::set_error("no HIP device");
        return DMX_E_HIP;
    }
    dmx::Engine* e = p->precision == DMX_FP32 ? dmx::make_engine_f32(*p, &rc) : dmx::make_engine_f64(*p, &rc);
    if (!e) return rc ? rc : DMX_E_HIP;
    *ctx = new dmx_ctx{e, *p};
    return DMX_OK;
}

void dmx_destroy(dmx_ctx* ctx)
{
    if (!ctx) return;
    delete ctx->e;
    delete ctx;
}

#define CTX_CHECK(c)                                              \
    do {                                                          \
        if (!(c) || !(c)->e) {                                    \
            dmx::set_error("null context");                       \
            return DMX_E_STATE;                                   \
        }                                                         \
    } while (0)

int dmx_upload(dmx_ctx* c, const double* uin)
{
    CTX_CHECK(c);
    if (!uin) { dmx::set_error("null uin"); return DMX_E_PARAM; }
    return c->e->upload(uin);
}
int dmx_download(dmx_ctx* c, double* uin)
{
    CTX_CHECK(c);
    if (!uin) { dmx::set_error("null uin"); return DMX_E_PARAM; }
    return c->e->download(uin);
}
int dmx_compute_dt(dmx_ctx* c, double* dt)
{
    CTX_CHECK(c);
    if (!dt) { dmx::set_error("null dt"); return DMX_E_PARAM; }
    return c->e->compute_dt(dt);
}
int dmx_compute_dt_noalloc(const dmx_params* p, const double* uin, double* dt, void* stream)
{
    int rc = validate(p);
    if (rc) return rc;
    if (p->precision != DMX_FP64) {
        dmx::set_error("dmx_compute_dt_noalloc reads a float64 state: precision must be DMX_FP64");
        return DMX_E_PARAM;
    }
    if (!uin || !dt) { dmx::set_error("null uin or dt"); return DMX_E_PARAM; }
    return dmx::noalloc_compute_dt_f64(*p, uin, dt, stream);
}
int dmx_godunov(dmx_ctx* c, double time, double dt)
{
    CTX_CHECK(c);
    if (!(dt > 0.0)) { dmx::set_error("dt must be > 0"); return DMX_E_PARAM; }
    return c->e->godunov(time, dt);
}
int dmx_fargo(dmx_ctx* c, double dt)
{
    CTX_CHECK(c);
    return c->e->fargo(dt);
}
int dmx_boundary(dmx_ctx* c, double time, double dt)
{
    CTX_CHECK(c);
    return c->e->boundary(time, dt);
}
int dmx_dissipation(dmx_ctx* c, double time, double dt)
{
    CTX_CHECK(c);
    if (!(dt > 0.0)) { dmx::set_error("dt must be > 0"); return DMX_E_PARAM; }
    return c->e->dissipation(time, dt);
}
int dmx_advance(dmx_ctx* c, int32_t nsteps, double* time, double* dts)
{
    CTX_CHECK(c);
    if (!time) { dmx::set_error("null time"); return DMX_E_PARAM; }
    return c->e->advance(nsteps, time, dts);
}
int dmx_history(dmx_ctx* c, double out[9])
{
    CTX_CHECK(c);
    if (!out) { dmx::set_error("null out"); return DMX_E_PARAM; }
    return c->e->history(0, out);
}
int dmx_history_shear(dmx_ctx* c, double out[9])
{
    CTX_CHECK(c);
    if (!out) { dmx::set_error("null out"); return DMX_E_PARAM; }
    return c->e->history(1, out);
}
int dmx_last_advance_ms(dmx_ctx* c, double* ms)
{
    CTX_CHECK(c);
    if (!ms) { dmx::set_error("null ms"); return DMX_E_PARAM; }
    return c->e->last_advance_ms(ms);
}
int dmx_set_limiter(dmx_ctx* c, const float* theta, int32_t n)
{
    CTX_CHECK(c);
    if (!theta) { dmx::set_error("null theta"); return DMX_E_PARAM; }
    return c->e->set_limiter(theta, n);
}
int dmx_learned_slopes(dmx_ctx* c, float* out, double* fallback)
{
    CTX_CHECK(c);
    if (!out) { dmx::set_error("null out"); return DMX_E_PARAM; }
    return c->e->learned_slopes(out, fallback);
}
int dmx_tune(dmx_ctx* c, double* best_ms, int32_t* shape)
{
    CTX_CHECK(c);
    int sh[4] = {0, 0, 0, 0};
    int rc = c->e->tune(best_ms, sh);
    if (!rc && shape)
        for (int q = 0; q < 4; q++) shape[q] = sh[q];
    return rc;
}
int dmx_set_shape(dmx_ctx* c, const int32_t* shape)
{
    CTX_CHECK(c);
    if (!shape) { dmx::set_error("null shape"); return DMX_E_PARAM; }
    int sh[4] = {shape[0], shape[1], shape[2], shape[3]};
    return c->e->set_shape(sh);
}
int dmx_profile(dmx_ctx* c, int32_t nsteps, double* time, double* ms, int32_t nkernels)
{
    CTX_CHECK(c);
    if (!time || !ms || nkernels < 0) { dmx::set_error("bad profile arguments"); return DMX_E_PARAM; }
    return c->e->profile(nsteps, time, ms, nkernels);
}
// ---------------------------------------------------------------------------
// The reference's binary dump (writer src/subroutines/output.f90:224-235,
// reader src/subroutines/restart.f90:212-223): a Fortran unformatted
// sequential file of five records, each framed by 4-byte little-endian
// length markers:
//   int32[6] nx, ny, nz, nxslice, nyslice, nzslice
//   f64[5]   time, dt, dx, dy, dz
//   f64[6]   xmin, xmax, ymin, ymax, zmin, zmax
//   f64[2]   gamma, ciso
//   f64      uin(-2:nx+3, -2:ny+3, -2:nz+3, 1:11)
// ---------------------------------------------------------------------------
static bool put_record(FILE* f, const void* data, size_t n)
{
    int32_t m = (int32_t)n;
    return fwrite(&m, 4, 1, f) == 1 && (n == 0 || fwrite(data, 1, n, f) == n) && fwrite(&m, 4, 1, f) == 1;
}

static bool get_record(FILE* f, void* data, size_t n)
{
    int32_t m0 = 0, m1 = 0;
    if (fread(&m0, 4, 1, f) != 1 || m0 < 0 || (size_t)m0 != n) return false;
    if (n && fread(data, 1, n, f) != n) return false;
    return fread(&m1, 4, 1, f) == 1 && m1 == m0;
}

static size_t uin_count(const dmx_params& p)
{
    return (size_t)(p.nx + 6) * (size_t)(p.ny + 6) * (size_t)(p.nz + 6) * 11;
}

int dmx_write_dump(dmx_ctx* c, const char* path, double time, double dt)
{
    CTX_CHECK(c);
    if (!path) { dmx::set_error("null path"); return DMX_E_PARAM; }
    const dmx_params& p = c->p;
    const size_t n = uin_count(p);
    if (n * 8 > 0x7fffffffull) {
        dmx::set_error("uin record of %zu bytes exceeds the 2 GiB record of the binary format", n * 8);
        return DMX_E_PARAM;
    }
    std::vector<double> u(n);
    int rc = c->e->download(u.data());
    if (rc) return rc;
    FILE* f = fopen(path, "wb");
    if (!f) { dmx::set_error("cannot open %s for writing", path); return DMX_E_PARAM; }
    const int32_t hdr[6] = {p.nx, p.ny, p.nz, 1, 1, 1};
    const double r2[5] = {time, dt, (p.xmax - p.xmin) / p.nx, (p.ymax - p.ymin) / p.ny, (p.zmax - p.zmin) / p.nz};
    const double r3[6] = {p.xmin, p.xmax, p.ymin, p.ymax, p.zmin, p.zmax};
    const double r4[2] = {p.gamma, p.ciso};
    bool ok = put_record(f, hdr, sizeof(hdr)) && put_record(f, r2, sizeof(r2)) && put_record(f, r3, sizeof(r3))
              && put_record(f, r4, sizeof(r4)) && put_record(f, u.data(), n * 8);
    ok = (fclose(f) == 0) && ok;
    if (!ok) { dmx::set_error("write error on %s", path); return DMX_E_PARAM; }
    return DMX_OK;
}

int dmx_read_dump(dmx_ctx* c, const char* path, double* time, double* dt)
{
    CTX_CHECK(c);
    if (!path || !time || !dt) { dmx::set_error("null argument"); return DMX_E_PARAM; }
    const dmx_params& p = c->p;
    FILE* f = fopen(path, "rb");
    if (!f) { dmx::set_error("cannot open %s", path); return DMX_E_PARAM; }
    int32_t hdr[6];
    double r2[5], r3[6], r4[2];
    bool ok = get_record(f, hdr, sizeof(hdr)) && get_record(f, r2, sizeof(r2)) && get_record(f, r3, sizeof(r3))
              && get_record(f, r4, sizeof(r4));
    if (!ok) { fclose(f); dmx::set_error("%s: not a binary dump (bad record framing)", path); return DMX_E_PARAM; }
    if (hdr[0] != p.nx || hdr[1] != p.ny || hdr[2] != p.nz || hdr[3] != 1 || hdr[4] != 1 || hdr[5] != 1) {
        fclose(f);
        dmx::set_error("%s: dump of %dx%dx%d on %dx%dx%d ranks does not match the %dx%dx%d context", path, hdr[0],
                       hdr[1], hdr[2], hdr[3], hdr[4], hdr[5], p.nx, p.ny, p.nz);
        return DMX_E_PARAM;
    }
    const double want3[6] = {p.xmin, p.xmax, p.ymin, p.ymax, p.zmin, p.zmax};
    if (memcmp(r3, want3, sizeof(r3)) != 0 || r4[0] != p.gamma || r4[1] != p.ciso) {
        fclose(f);
        dmx::set_error("%s: domain bounds, gamma or ciso differ from the context's", path);
        return DMX_E_PARAM;
    }
    const size_t n = uin_count(p);
    std::vector<double> u(n);
    ok = get_record(f, u.data(), n * 8);
    fclose(f);
    if (!ok) { dmx::set_error("%s: truncated or mismatched uin record", path); return DMX_E_PARAM; }
    int rc = c->e->upload(u.data());
    if (rc) return rc;
    *time = r2[0];
    *dt = r2[1];
    return DMX_OK;
}

struct dmx_batch {
    dmx::Batch* b;
    int32_t n;
};

int dmx_batch_create(dmx_ctx* const* ctxs, int32_t n, int32_t nstreams, dmx_batch** batch)
{
    if (!batch) { dmx::set_error("null batch"); return DMX_E_PARAM; }
    *batch = nullptr;
    if (!ctxs || n < 1) { dmx::set_error("batch: no contexts"); return DMX_E_PARAM; }
    std::vector<dmx::Engine*> es(n);
    for (int32_t b = 0; b < n; b++) {
        if (!ctxs[b] || !ctxs[b]->e) { dmx::set_error("batch: null context %d", b); return DMX_E_STATE; }
        for (int32_t a = 0; a < b; a++)
            if (ctxs[a] == ctxs[b]) { dmx::set_error("batch: context %d repeats context %d", b, a); return DMX_E_PARAM; }
        es[b] = ctxs[b]->e;
    }
    int rc = 0;
    dmx::Batch* bt = ctxs[0]->p.precision == DMX_FP32 ? dmx::make_batch_f32(es.data(), n, nstreams, &rc)
                                                      : dmx::make_batch_f64(es.data(), n, nstreams, &rc);
    if (!bt) return rc ? rc : DMX_E_HIP;
    *batch = new dmx_batch{bt, n};
    return DMX_OK;
}

void dmx_batch_destroy(dmx_batch* batch)
{
    if (!batch) return;
    delete batch->b;
    delete batch;
}

int dmx_batch_advance(dmx_batch* batch, int32_t nsteps, double* times)
{
    if (!batch || !batch->b) { dmx::set_error("null batch"); return DMX_E_STATE; }
    if (!times) { dmx::set_error("null times"); return DMX_E_PARAM; }
    return batch->b->advance(nsteps, times);
}

int dmx_batch_tune(dmx_batch* batch, double* best_ms, int32_t* shape)
{
    if (!batch || !batch->b) { dmx::set_error("null batch"); return DMX_E_STATE; }
    int sh[4] = {0, 0, 0, 0};
    int rc = batch->b->tune(best_ms, sh);
    if (!rc && shape)
        for (int q = 0; q < 4; q++) shape[q] = sh[q];
    return rc;
}

int dmx_batch_last_ms(dmx_batch* batch, double* ms)
{
    if (!batch || !batch->b) { dmx::set_error("null batch"); return DMX_E_STATE; }
    if (!ms) { dmx::set_error("null ms"); return DMX_E_PARAM; }
    return batch->b->last_ms(ms);
}

const char* dmx_kernel_name(int32_t id)
{
    return (id >= 0 && id < dmx::KG_N) ? dmx::kKernelGroupNames[id] : "";
}
const char* dmx_last_error(void) { return dmx::g_err; }
const char* dmx_version(void) { return "dmx 0.1 (gfx950)"; }

} // extern "C"
