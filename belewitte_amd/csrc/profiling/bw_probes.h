/*
 * bw_probes.h -- measurement probes of libbw_probes.so.
 *
 * Not part of the kernel library's ABI (include/belewitte_hip.h): these
 * kernels exist only to calibrate and bound what the profiles report, and
 * belewitte_amd/profiling.py is their only caller.  Status convention as
 * libbelewitte_hip: 0 on success, non-zero with bw_probe_last_error().
 */
#ifndef BW_PROBES_H
#define BW_PROBES_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

const char *bw_probe_last_error(void);

/* Streaming copy of n 32-bit words (dst[i] = src[i]), 16-byte aligned
 * buffers: reads and writes exactly 4n bytes, the reference point that
 * scales the FETCH_SIZE / WRITE_SIZE counters to bytes. */
int bw_probe_copy(uint32_t *dst_device, const uint32_t *src_device, uint64_t n, void *stream);

/* Gather-rate probe: out[t] = sum over r of table[idx[r * threads + t]]
 * (indices >= table_len read 0), 16 loads in flight per thread, so the
 * lanes of one wave read consecutive index entries.  n must be a positive
 * multiple of 16 * threads, threads a multiple of 256, out holds `threads`
 * floats.  The launch time goes to *kernel_ms when it is not NULL. */
int bw_probe_gather(const uint32_t *idx_device, uint64_t n, const float *table_device, uint32_t table_len,
                    float *out_device, uint32_t threads, float *kernel_ms, void *stream);

#ifdef __cplusplus
}
#endif
#endif
