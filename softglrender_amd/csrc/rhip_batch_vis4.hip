/*
 * rhip_batch_vis4.hip -- the persistent raster of batched views, four
 * samples: tile edges 8 to 32 pixels with the wavefront counts that give each
 * lane 1 to 4 pixels (16 sample keys), for every depth test it takes.
 */
#include "rhip_batch.h"

namespace rhip {

static const VisKernel kVis4[] = {
    RHIP_VIS_MODES(4, 8, 1),  RHIP_VIS_MODES(4, 16, 1), RHIP_VIS_MODES(4, 16, 2), RHIP_VIS_MODES(4, 16, 4),
    RHIP_VIS_MODES(4, 32, 4), RHIP_VIS_MODES(4, 32, 8), RHIP_VIS_MODES(4, 32, 16),
};

const VisKernel *visKernels4(int *n) {
  *n = (int)(sizeof kVis4 / sizeof kVis4[0]);
  return kVis4;
}

}  // namespace rhip
