// One storage scalar's instantiation of the engine drivers. build.py compiles
// this file four times (-DVMC_ENGINE_T=float|double|cf32|cf64 with the matching
// -DVMC_ENGINE_OPS name), and the C ABI in engine.hip dispatches on the plan
// dtype to the exported table.
#include "engine_core.hpp"

#ifndef VMC_ENGINE_T
#error "VMC_ENGINE_T (storage scalar) must be defined"
#endif
#ifndef VMC_ENGINE_OPS
#error "VMC_ENGINE_OPS (exported table name) must be defined"
#endif

namespace vmc {
namespace {

using T = VMC_ENGINE_T;

int amp_grad(vmc_plan *p, const void *t, const int32_t *cfg, int batch, double *la, void *o, int8_t *pp, void *ws,
             size_t wsb, hipStream_t st) {
  return amp_grad_impl<T>(p, (const T *)t, cfg, batch, la, (T *)o, pp, ws, wsb, st);
}

int init_cache(vmc_plan *p, const void *t, const int32_t *cfg, int n, void *cache, void *ws, size_t wsb,
               hipStream_t st) {
  return init_cache_impl<T>(p, (const T *)t, cfg, n, cache, ws, wsb, st);
}

int sweep(vmc_plan *p, const void *t, int32_t *cfg, int n, uint64_t seed, int sw, void *cache, double *la, void *o,
          int8_t *pp, double *eloc, void *ws, size_t wsb, hipStream_t st, bool tr, bool es) {
  return sweep_impl<T>(p, (const T *)t, cfg, n, seed, sw, cache, la, (T *)o, pp, eloc, ws, wsb, st, tr, es);
}

int top(vmc_plan *p, const void *t, const int32_t *cfg, double *la, int n, void *ws, hipStream_t st) {
  const Layout L = make_layout(p, n, p->has_v);
  top_sweep<T>(p, bufs<T>(L, ws), (const T *)t, cfg, la, n, st);
  return 0;
}

int gi_amp_grad(vmc_plan *p, const void *t, const int32_t *s, int batch, double *la, void *o, int8_t *pp, void *ws,
                size_t wsb, hipStream_t st) {
  return gi_amp_grad_impl<T>(p, (const T *)t, s, batch, la, (T *)o, pp, ws, wsb, st);
}

int gi_init_cache(vmc_plan *p, const void *t, const int32_t *s, int n, void *cache, void *ws, size_t wsb,
                  hipStream_t st) {
  return gi_init_cache_impl<T>(p, (const T *)t, s, n, cache, ws, wsb, st);
}

int gi_sweep(vmc_plan *p, const void *t, int32_t *s, int n, uint64_t seed, int sw, void *cache, double *la, void *o,
             int8_t *pp, double *eloc, void *ws, size_t wsb, hipStream_t st) {
  return gi_sweep_impl<T>(p, (const T *)t, s, n, seed, sw, cache, la, (T *)o, pp, eloc, ws, wsb, st);
}

int bl_sweep(vmc_plan *p, const void *t, int32_t *s, int n, uint64_t seed, int sw, void *cache, double *la, void *o,
             int8_t *pp, double *eloc, void *ws, size_t wsb, hipStream_t st) {
  return bl_sweep_impl<T>(p, (const T *)t, s, n, seed, sw, cache, la, (T *)o, pp, eloc, ws, wsb, st);
}

}  // namespace

EngineOps VMC_ENGINE_OPS = {amp_grad, init_cache, sweep, top, gi_amp_grad, gi_init_cache, gi_sweep, bl_sweep};

}  // namespace vmc
