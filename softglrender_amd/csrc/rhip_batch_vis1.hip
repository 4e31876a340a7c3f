/*
 * rhip_batch_vis1.hip -- the persistent raster of batched views, single
 * sample: every tile edge (8, 16, 32, 64 pixels) with the wavefront counts
 * that give each lane 1 to 16 pixels, for every depth test it takes.
 */
#include "rhip_batch.h"

namespace rhip {

static const VisKernel kVis1[] = {
    RHIP_VIS_MODES(1, 8, 1),  RHIP_VIS_MODES(1, 16, 1), RHIP_VIS_MODES(1, 16, 2), RHIP_VIS_MODES(1, 16, 4),
    RHIP_VIS_MODES(1, 32, 1), RHIP_VIS_MODES(1, 32, 2), RHIP_VIS_MODES(1, 32, 4), RHIP_VIS_MODES(1, 32, 8),
    RHIP_VIS_MODES(1, 32, 16), RHIP_VIS_MODES(1, 64, 4), RHIP_VIS_MODES(1, 64, 8), RHIP_VIS_MODES(1, 64, 16),
};

const VisKernel *visKernels1(int *n) {
  *n = (int)(sizeof kVis1 / sizeof kVis1[0]);
  return kVis1;
}

}  // namespace rhip
