// float64 engine: the reference's arithmetic, compiled with -ffp-contract=off
// so every kernel reproduces the reference bit for bit.
#define DMX_T double
#define DMX_NS f64
#include "dmx_step.inc"

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
