/* PreCorr device parameters; the per-channel state layout is the seam's
 * (gpu_precorr_state_t, include/gpu_compute.h:128-136). */
#ifndef PC_TYPES_H
#define PC_TYPES_H
#include <stdint.h>
#include "../../include/dsd_hip.h"

typedef struct pc_params {
    float a[8], g[8];
    float state_limit;
    int32_t order;
} pc_params;

#endif
