/*
 * libismc_diag — measurement probes for MI355X (gfx950), built beside
 * libismc but not part of it.  They have no counterpart in the reference:
 * they measure the machine that the transport kernel's roofline and the
 * byte counters of its profiles are read against.  Every entry point
 * returns 0 or a negative status and never exits; the last failure's
 * message is in ismc_diag_last_error (per host thread).
 */
#ifndef ISMC_DIAG_H
#define ISMC_DIAG_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ISMC_DIAG_EINVAL -1
#define ISMC_DIAG_EDEVICE -2
#define ISMC_DIAG_ENOMEM -3

/* VALU issue costs.  Op k (0 .. ismc_issue_rate_ops() - 1, named by
 * ismc_issue_rate_name) runs one instruction kind in 8 independent chains
 * per lane, 64 workgroups of 256 lanes per CU, five launches after a
 * warm-up; returns wave64 instructions per second and the shader clock the
 * loops ran at (s_memtime over the 100 MHz real-time counter).
 * cycles per instruction = 4 x CUs x clock / rate. */
int32_t ismc_issue_rate_ops(void);
const char* ismc_issue_rate_name(int32_t op);
int ismc_issue_rate(int device, int32_t op, double* wave_insts_per_s, double* shader_ghz);

/* Copy `bytes` bytes (a positive multiple of 16) between two device
 * buffers with a plain streaming kernel (copy_probe_kernel) and return its
 * device time: a known-size reference for the hardware byte counters
 * (FETCH_SIZE / WRITE_SIZE) of a profile run. */
int ismc_copy_probe(int device, int64_t bytes, float* kernel_ms);

const char* ismc_diag_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
