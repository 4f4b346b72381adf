// Reduced-precision assignment (placeholder until the MFMA kernels land).
#include "km_engine.hpp"
#include "km_mfma.hpp"

namespace km {
void mfma_free(Engine *) {}
void mfma_invalidate(Engine *) {}
hipError_t mfma_assign(Engine *, long, long) { return hipErrorNotSupported; }
hipError_t mfma_reresolve(Engine *, long, long) { return hipErrorNotSupported; }
hipError_t mfma_centers(Engine *) { return hipErrorNotSupported; }
void mfma_collect(Engine *) {}
}  // namespace km
