/* Internal declarations shared by the host-side translation units. */
#ifndef DSDHIP_INTERNAL_H
#define DSDHIP_INTERNAL_H

#include <stddef.h>
#include <stdint.h>
#include "tr_types.h"
#include "../../include/dsd_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Selects the device on first use; 0 on success. */
int dsdhip_ensure_device(void);

#ifdef __cplusplus
}
#endif
#endif
