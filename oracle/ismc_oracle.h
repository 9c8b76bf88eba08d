/*
 * ORACLE — CPU restatement of the reference ISMC/ISMLMC photon transport.
 *
 * This is test infrastructure.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it, and only as the checker.  The
 * product (ismc_amd/) never links, imports or calls anything in oracle/.
 *
 * Algorithm followed (reference files, cited as path:line):
 *   - MLMC.py:76-202       multilayer substep propagator (hop/drop/spin,
 *                           boundary substeps, Fresnel reflect/refract,
 *                           roulette 1e-4 / 1-in-10)
 *   - ISMLMC.py:79-97      glass entrance reflection and sphere return draw
 *   - ISMLMC.py:101-347    packet loop with dual-sphere capture, crosstalk
 *                           and re-incidence
 *   - ISMC.py:87-160       single-slab mirror propagator (flag ORACLE_MIRROR)
 *   - CUDA/CUDAISMLMC.h:71-129,132-314,316-654  (same semantics in CUDA;
 *                           FP32 arithmetic is selected with ORACLE_FP32)
 *
 * Random numbers: the reference draws from cuRAND XORWOW / Python MT19937.
 * The oracle (like the GPU path) draws from Philox4x32-10 (Salmon et al.,
 * SC'11) with key = seed and counter = (packet id lo, event index, packet
 * id hi, stream id); every event of a packet gets one 4-word block and the
 * reference's draw sites map onto fixed words of that block (see
 * ismc_oracle.c).  Parity with the reference is therefore statistical, and
 * parity between this oracle and the GPU kernel is per packet.
 *
 * Pinning of each function (what checks it, in tests/):
 *   philox      Random123 known-answer vectors + rocRAND's own Philox rounds
 *   fresnel     textbook closed forms (normal incidence, TIR, Brewster,
 *               reciprocity, grazing); snell: unit norm + Snell's invariant
 *   entrance    closed-form glass-entrance reflectance (0.0441282 for air/1.52 glass/1.40 sample at normal incidence)
 *   HG spin     E[cos]=g, E[P2]=g^2, KS vs HG CDF, |mu|=1
 *   sphere      closed-form return probability, black-sphere == no-sphere
 *   transport   absorption-only slab closed forms, matched-boundary ballistic
 *               T_u = exp(-mu_t d), van de Hulst / MCML known answer, energy
 *               conservation, mirror == substep per packet, and golden
 *               fixtures written by importing the reference's own Python
 *               propagators (tests/golden/make_golden.py)
 */
#ifndef ISMC_ORACLE_H
#define ISMC_ORACLE_H
#include <stdint.h>

#define ORACLE_MAX_LAYERS 16

#define ORACLE_FP32      1u   /* do the transport arithmetic in float       */
#define ORACLE_MIRROR    2u   /* ISMC.py single-slab mirror propagator       */
#define ORACLE_CUDA_ROULETTE 4u /* CUDAISMLMC.h:275 integer 1/m == 0: never survive */
#define ORACLE_ETALON    8u   /* a slide's bounces in closed form (one event);
                                 the default restates the reference's
                                 bounce-by-bounce simulation */

typedef struct {
    int32_t present;
    int32_t _pad;
    double pw, fs, fp, fd, f, angle;   /* angle in radians */
} oracle_sphere;

typedef struct {
    int32_t n_layers;
    int32_t specular_included;
    double n_ambient;
    double roulette_threshold;
    double roulette_m;
    double n[ORACLE_MAX_LAYERS];
    double g[ORACLE_MAX_LAYERS];
    double t[ORACLE_MAX_LAYERS];
    double mu_a[ORACLE_MAX_LAYERS];
    double mu_s[ORACLE_MAX_LAYERS];    /* mu_s == 0 marks a glass layer */
    double mu_t[ORACLE_MAX_LAYERS];
    oracle_sphere R, T;
} oracle_config;

/* tally index order: A, R_d, R_s, T_d, T_u  (ISMLMC.py:347) */
typedef struct {
    double sum[5];
    double sumsq[5];
    uint64_t events[5];   /* interactions, boundary hits, entrances,
                             sphere decisions, re-injections */
    uint64_t n_packets;
    uint64_t n_aborted;
} oracle_tally;

void oracle_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
double oracle_fresnel(double ni, double nt, double mu_z);
float oracle_fresnel_f(float ni, float nt, float mu_z);
double oracle_entrance_reflectance(const oracle_config* c, double mu_z);
/* the exit index that bounds the direction inside glass layer g (the glass-entry guard) */
double oracle_glass_guard_index(const oracle_config* c, int g);
double oracle_hg_cos(double g, double u);

/* per_packet (nullable): n_packets rows of {A, R_d, R_s, T_d, T_u, events} */
int oracle_simulate(const oracle_config* cfg, uint64_t n_packets, uint64_t seed,
                    uint64_t first_pid, uint32_t stream, uint32_t flags,
                    oracle_tally* out, double* per_packet);

/* same, split over OpenMP threads (per-packet results are identical) */
int oracle_simulate_omp(const oracle_config* cfg, uint64_t n_packets, uint64_t seed,
                        uint64_t first_pid, uint32_t stream, uint32_t flags,
                        oracle_tally* out, int n_threads);
#endif
