// float32-bulk engine: hydro state, trace, Riemann solvers and fluxes in
// float32; face B, constrained-transport accumulation, pressure recovery and
// the history sums in float64.
#define DMX_T float
#define DMX_NS f32
#include "dmx_step.inc"

namespace dmx {
Engine* make_engine_f32(const dmx_params& p, int* rc)
{
    f32::EngineT* e = new f32::EngineT();
    *rc = e->init(p);
    if (*rc) {
        delete e;
        return nullptr;
    }
    return e;
}
Batch* make_batch_f32(Engine* const* es, int n, int nstreams, int* rc)
{
    f32::BatchT* b = new f32::BatchT();
    *rc = b->init(es, n, nstreams);
    if (*rc) {
        delete b;
        return nullptr;
    }
    return b;
}
} // namespace dmx
