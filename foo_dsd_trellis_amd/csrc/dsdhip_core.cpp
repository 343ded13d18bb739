/*
 * Device selection, NTF table and the gpu_compute.h backend seam
 * (include/gpu_compute.h:22-59, 185-233) for the HIP backend.
 */
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <mutex>
#include "dsdhip_internal.h"

/* CRFB coefficient table, same rows and order as the reference's
 * (src/ntf.c:16-792). */
static const dsdhip_ntf_t g_ntf[] = {
#include "ntf_table.inc"
};

static std::once_flag g_dev_once;
static int g_dev_state = -1;

extern "C" {

int dsdhip_ensure_device(void)
{
    std::call_once(g_dev_once, []() {
        int n = 0;
        if (hipGetDeviceCount(&n) != hipSuccess || n < 1) { g_dev_state = -1; return; }
        g_dev_state = 0;
    });
    return g_dev_state;
}

const char *dsdhip_version(void) { return "dsdhip 0.1 gfx950"; }

int dsdhip_ntf_count(void) { return (int)(sizeof(g_ntf) / sizeof(g_ntf[0])); }

const dsdhip_ntf_t *dsdhip_ntf_get(int index)
{
    if (index < 0 || index >= dsdhip_ntf_count()) return NULL;
    return &g_ntf[index];
}

/* ids as include/dsd_types.h:57-68 */
static const char *const k_names[10] = {"clans-4", "sdm-4", "clans-5", "sdm-5", "clans-6",
                                        "sdm-6", "clans-7", "sdm-7", "clans-8", "sdm-8"};

int dsdhip_ntf_lookup(int id, unsigned rate)
{
    if (id < 0 || id >= 10) return -1;
    for (int i = 0; i < dsdhip_ntf_count(); i++)
        if (strcmp(g_ntf[i].name, k_names[id]) == 0 && g_ntf[i].freq <= rate) return i;
    return -1;
}

int dsdhip_ntf_auto(unsigned rate)
{
    int id = rate >= 11289600u ? 8 : rate >= 5644800u ? 4 : 2;   /* src/ntf.c:839-855 */
    return dsdhip_ntf_lookup(id, rate);
}

bool gpu_hip_probe(void) { return dsdhip_ensure_device() == 0; }

void gpu_hip_get_info(gpu_hip_info_t *info)
{
    if (!info) return;
    memset(info, 0, sizeof(*info));
    info->backend = GPU_BACKEND_HIP;
    if (dsdhip_ensure_device() != 0) return;
    hipDeviceProp_t p;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&p, dev) != hipSuccess) return;
    snprintf(info->device_name, sizeof(info->device_name), "%s (%s)", p.name, p.gcnArchName);
    info->vram_mb = p.totalGlobalMem >> 20;
    info->available = true;
}

} /* extern "C" */
