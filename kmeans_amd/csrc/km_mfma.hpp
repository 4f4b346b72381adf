// Reduced-precision (BF16/FP16 MFMA) assignment with FP32 re-resolve.
#pragma once
#include "km_common.hpp"

namespace km {
struct Engine;
void mfma_free(Engine *e);
void mfma_invalidate(Engine *e);
void mfma_bounds_invalidate(Engine *e);  // labels changed outside mfma_assign
hipError_t mfma_assign(Engine *e, long r0, long rows);
// also makes the engine's stream wait for the two-candidate decisions that
// mfma_assign left running on its side stream
hipError_t mfma_reresolve(Engine *e, long r0, long rows);
// the engine's stream waits for those decisions (no-op when none are pending)
hipError_t mfma_pair_join(Engine *e);
hipError_t mfma_centers(Engine *e);
void mfma_collect(Engine *e, int nslots);
void mfma_centers_changed(Engine *e, bool new_range);
hipError_t mfma_ensure_slots(Engine *e, int n);
hipError_t mfma_prepare(Engine *e);
hipError_t mfma_prepare_rows(Engine *e, const float *X, long rows);
// xn: per point upper bounds of |x| and |x - xh/s| (FP32 units)
// order, home (nullable, both or neither): the label-ordered variant over all
// resident points (order = the previous update's sort by label, or the points
// the drift bounds left open, Mdev of them); xbo: fresh drift bounds
hipError_t mfma_assign_view(Engine *e, const float *X, const uint16_t *Xh, const float2 *xn,
                            int *lab, long rows, const int *order = nullptr,
                            const int *home = nullptr, const int *Mdev = nullptr,
                            float4 *xbo = nullptr, bool zero_counters = true);
hipError_t mfma_reresolve_view(Engine *e, const float *X, const uint16_t *Xh, const float2 *xn,
                               int *lab, long rows, float4 *xbo = nullptr);
const uint16_t *mfma_tile_xh(Engine *e);
const float2 *mfma_tile_xnorm(Engine *e);
int mfma_row_stride(Engine *e);  // padded width Dp of the reduced-precision rows
// Read-only view of the reduced-precision state (valid after mfma_prepare).
struct MfmaView {
  const uint16_t *Xh, *Ch;
  const float *bias, *bnd;    // bnd = [max |c_j|, max |c_j|^2, max |c_j - ch_j/s|] bounds
  const float2 *xnorm;        // per point bounds of |x| and |x - xh/s|
  int Dp, KS, Kp;
  bool fp16;
};
bool mfma_view(Engine *e, MfmaView *v);
bool mfma_ready(Engine *e);      // reduced-precision copy of every point present
}  // namespace km
