// Learned candidate proposer for the assignment step.
//
// The K centres are partitioned once into G groups (a coarse K-means of the
// centres themselves).  A proposer names T groups per point: either the MLP
// (din -> hidden -> hidden -> G, trained on group(label) of points sampled
// during the first plain iterations) or the T nearest group centroids.  Exact
// canonical FP32 scores are computed only against the members of the named
// groups; the best candidate's label is then certified by one of two
// triangle-inequality bounds (centre separation, or the group bound
// ||x - c|| >= ||x - G_g|| - r_g for every group not named), each widened by
// the FP32 rounding bound of the canonical score.  Uncertified points are
// searched in full by the plain FP32 kernel, so the labels are always the
// plain FP32 labels.
#pragma once
#include <vector>

#include "km_common.hpp"

namespace km {
struct Engine;

// Option defaults and validation (called from Engine::create).
kmeansErrorStatus learned_configure(Engine *e);
// Sampling of training rows after the assignment of plain iteration `it`.
hipError_t learned_collect(Engine *e, int it);
// Group construction and (MLP proposer) training; idempotent.
kmeansErrorStatus learned_prepare(Engine *e);
// Proposer assignment: proposal, grouped exact scores, certification.
hipError_t learned_assign(Engine *e);
// Full FP32 search of the points no bound certified.
hipError_t learned_fallback(Engine *e);
// Copies the device counters into e->stats.
void learned_stats(Engine *e);
void learned_free(Engine *e);
bool learned_active(const Engine *e);  // proposer iterations have started
}  // namespace km
