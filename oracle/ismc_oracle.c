/*
 * ORACLE — CPU restatement of the reference ISMC/ISMLMC transport.
 * Test infrastructure only; see ismc_oracle.h for what follows which
 * reference line and what pins each function.
 */
#define _GNU_SOURCE
#include "ismc_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* Philox4x32-10, Salmon, Moraes, Dror & Shaw, "Parallel random numbers:
 * as easy as 1, 2, 3", SC'11: round multipliers and Weyl key increments as
 * published there. */
void oracle_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int round = 0; round < 10; ++round) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        c0 = hi1 ^ c1 ^ k0;
        c1 = lo1;
        c2 = hi0 ^ c3 ^ k1;
        c3 = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0;
    out[1] = c1;
    out[2] = c2;
    out[3] = c3;
}

/* ---- double instantiation ---- */
#define REAL double
#define SFX _d
#define RABS fabs
#define RACOS acos
#define RASIN asin
#define RSIN sin
#define RCOS cos
#define RTAN tan
#define RSQRT sqrt
#define RLOG log
#define RMAX fmax
#define RCOPYSIGN copysign
#include "ismc_oracle_body.inc"
#undef REAL
#undef SFX
#undef RABS
#undef RACOS
#undef RASIN
#undef RSIN
#undef RCOS
#undef RTAN
#undef RSQRT
#undef RLOG
#undef RMAX
#undef RCOPYSIGN

/* ---- float instantiation ---- */
#define REAL float
#define SFX _f
#define RABS fabsf
#define RACOS acosf
#define RASIN asinf
#define RSIN sinf
#define RCOS cosf
#define RTAN tanf
#define RSQRT sqrtf
#define RLOG logf
#define RMAX fmaxf
#define RCOPYSIGN copysignf
#include "ismc_oracle_body.inc"

double oracle_fresnel(double ni, double nt, double mu_z) { return fresnel_d(ni, nt, mu_z); }
float oracle_fresnel_f(float ni, float nt, float mu_z) { return fresnel_f(ni, nt, mu_z); }

double oracle_entrance_reflectance(const oracle_config* c, double mu_z)
{
    int layer = mu_z > 0 ? 0 : c->n_layers - 1;
    return entrance_R_d(c, layer, mu_z);
}

double oracle_glass_guard_index(const oracle_config* c, int g) { return glass_guard_index_d(c, g); }

double oracle_hg_cos(double g, double u)
{
    if (g == 0.0) return 2 * u - 1;
    double q = (1 - g * g) / (1 - g + 2 * g * u);
    return (1 / (2 * g)) * (1 + g * g - q * q);
}

int oracle_simulate(const oracle_config* cfg, uint64_t n_packets, uint64_t seed,
                    uint64_t first_pid, uint32_t stream, uint32_t flags, oracle_tally* out,
                    double* per_packet)
{
    if (!cfg || !out || cfg->n_layers < 1 || cfg->n_layers > ORACLE_MAX_LAYERS) return -1;
    if ((flags & ORACLE_MIRROR) && cfg->n_layers != 1) return -1;
    if (flags & ORACLE_FP32)
        return simulate_f(cfg, n_packets, seed, first_pid, stream, flags, out, per_packet);
    return simulate_d(cfg, n_packets, seed, first_pid, stream, flags, out, per_packet);
}

int oracle_simulate_omp(const oracle_config* cfg, uint64_t n_packets, uint64_t seed,
                        uint64_t first_pid, uint32_t stream, uint32_t flags, oracle_tally* out,
                        int n_threads)
{
    if (!cfg || !out || n_threads < 1) return -1;
    oracle_tally* parts = (oracle_tally*)calloc((size_t)n_threads, sizeof(oracle_tally));
    if (!parts) return -2;
    int rc = 0;
#ifdef _OPENMP
#pragma omp parallel for num_threads(n_threads) schedule(static, 1) reduction(| : rc)
#endif
    for (int t = 0; t < n_threads; ++t) {
        uint64_t lo = n_packets * (uint64_t)t / (uint64_t)n_threads;
        uint64_t hi = n_packets * (uint64_t)(t + 1) / (uint64_t)n_threads;
        rc |= oracle_simulate(cfg, hi - lo, seed, first_pid + lo, stream, flags, &parts[t], NULL);
    }
    memset(out, 0, sizeof(*out));
    for (int t = 0; t < n_threads; ++t) {
        for (int i = 0; i < 5; ++i) {
            out->sum[i] += parts[t].sum[i];
            out->sumsq[i] += parts[t].sumsq[i];
            out->events[i] += parts[t].events[i];
        }
        out->n_packets += parts[t].n_packets;
        out->n_aborted += parts[t].n_aborted;
    }
    free(parts);
    return rc;
}
