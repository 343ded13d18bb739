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

int dsdhip_trellis_get(dsdhip_trellis_t *t, int i, tr_ctx *out);
int dsdhip_trellis_put(dsdhip_trellis_t *t, int i, const tr_ctx *in);
int dsdhip_trellis_clone(dsdhip_trellis_t *t, int dst, int src);
int dsdhip_trellis_detach(dsdhip_trellis_t *t, int i);
int dsdhip_trellis_copy_detached(dsdhip_trellis_t *t, int dst, int src);
int dsdhip_trellis_params(const dsdhip_trellis_t *t, tr_params *out);
void *dsdhip_trellis_context_ptr(dsdhip_trellis_t *t, int i);
/* 255-tap Kaiser low-pass of the stitch quality (src/dsp_plugin.c:2571-2604). */
void dsdhip_make_qfir(double *q);
/* Device scratch of the FIR chain for n input samples per channel, so later
 * dsdhip_fir_process_dev calls of at most n samples allocate nothing. */
int dsdhip_fir_reserve(dsdhip_fir_t *f, size_t n);
/* f64 -> f32 copy on the device (round to nearest). */
int dsdhip_f64_to_f32_dev(const double *d_x, float *d_y, size_t n, void *stream);

#ifdef __cplusplus
}
#endif
#endif
