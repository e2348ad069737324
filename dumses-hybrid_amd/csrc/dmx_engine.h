// dmx_engine.h -- host-side engine interface behind the C ABI.
//
// One Engine owns one device-resident domain and one HIP stream.  The float64
// and float32 engines are built from the same kernel source in two
// translation units (dmx_f64.hip, dmx_f32.hip); dmx_abi.cpp picks one per
// context from dmx_params.precision.
#pragma once

#include <stdint.h>

#include "dmx.h"

namespace dmx {

class Engine {
public:
    virtual ~Engine() {}
    virtual int upload(const double* uin) = 0;
    virtual int download(double* uin) = 0;
    virtual int compute_dt(double* dt) = 0;
    virtual int godunov(double time, double dt) = 0;
    virtual int fargo(double dt) = 0;
    virtual int boundary(double time, double dt) = 0;
    virtual int dissipation(double time, double dt) = 0;
    virtual int advance(int nsteps, double* time, double* dts) = 0;
    virtual int history(int variant, double out[9]) = 0;
    virtual int last_advance_ms(double* ms) = 0;
    virtual int profile(int nsteps, double* time, double* ms, int nk) = 0;
    virtual int tune(double* best_ms, int* shape) = 0;
    virtual int set_shape(const int* shape) = 0;
    virtual int set_limiter(const float* theta, int n) = 0;
    virtual int learned_slopes(float* out, double* fallback) = 0;
};

// Several contexts of one precision advanced together (dmx_batch_*): every
// kernel of the step is one launch over the tiles of all boxes.
class Batch {
public:
    virtual ~Batch() {}
    virtual int advance(int nsteps, double* times) = 0;
    virtual int tune(double* best_ms, int* shape) = 0;
    virtual int last_ms(double* ms) = 0;
};

// kernel groups timed by dmx_profile
enum KernelGroup { KG_DT = 0, KG_TRACE, KG_FACES, KG_EDGES, KG_SHEARFLUX, KG_UPDATE, KG_FARGO, KG_BOUNDARY,
                   KG_FLUX, KG_DISSIPATION, KG_LEARNED, KG_FUSED, KG_N };
static const char* const kKernelGroupNames[KG_N] = {
    "compute_dt", "trace", "faces", "edges", "shearing_flux", "update_ct", "fargo", "boundary", "flux",
    "dissipation", "learned_limiter", "fused"};

// set by the engines, read by dmx_last_error()
void set_error(const char* fmt, ...);

Engine* make_engine_f64(const dmx_params& p, int* rc);
Engine* make_engine_f32(const dmx_params& p, int* rc);
Batch* make_batch_f64(Engine* const* es, int n, int nstreams, int* rc);
Batch* make_batch_f32(Engine* const* es, int n, int nstreams, int* rc);
// dmx_compute_dt_noalloc (float64 state in the caller's device memory;
// stream is a hipStream_t)
int noalloc_compute_dt_f64(const dmx_params& p, const double* uin, double* dt, void* stream);

} // namespace dmx
