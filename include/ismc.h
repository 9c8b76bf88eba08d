/*
 * libismc — MI355X (gfx950) integrating-sphere Monte Carlo photon transport.
 *
 * C ABI (plain pointers and sizes) loaded by ctypes from ismc_amd/.  Every
 * entry point returns 0 on success or a negative status; it never exits
 * the process.  The caller owns all host buffers; the context owns device
 * memory.  One context per host thread; calls block until their results
 * are on the host.
 *
 * Reference interfaces replaced (pdcook/ISMC, cited path:line):
 *   - ismc_simulate        <- ISMLMC() host wrapper, CUDA/CUDAISMLMC.h:656-755,
 *                             and the Python propagators ISMC(), ISMC.py:164-406,
 *                             ISMLMC(), ISMLMC.py:101-347
 *   - ismc_simulate_batch  <- the per-point launch loop of
 *                             singleLayerSampleSpace, CUDA/ISMLMC-lookup.cu:153-421,
 *                             and the per-evaluation launches of cost(),
 *                             CUDA/ISMLMC-minimization.cu:89-131
 *   - ismc_create/destroy  <- the process-global RNG state set-up
 *                             (rand_set/globalDeviceStates, CUDA/CUDAISMLMC.h:19-25,
 *                             670-683) and cudaSetDevice (:660)
 *   - ismc_device_info     <- CUDABasicProperties, CUDA/CUDAISMLMC.h:28-38
 *   - ismc_last_error      <- gpuAssert's message, CUDA/CUDAISMLMC.h:41-48
 */
#ifndef ISMC_H
#define ISMC_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ISMC_MAX_LAYERS 16

/* status codes */
#define ISMC_OK 0
#define ISMC_EINVAL -1      /* invalid argument or configuration          */
#define ISMC_EDEVICE -2     /* HIP runtime error (message in last_error)   */
#define ISMC_ENOMEM -3      /* device or pinned host allocation failed     */
#define ISMC_ERANGE -4      /* a size exceeds what one launch can tally    */

/* flags */
#define ISMC_FLAG_FP64 0x1u           /* FP64 transport (deviation reference)  */
#define ISMC_FLAG_CUDA_ROULETTE 0x4u  /* CUDAISMLMC.h:275: int 1/m == 0, no survivor */
/* Simulate a slide's bounces face by face, as MLMC.py:114-163 does.  By
 * default a slide (a glass layer between scattering layers or the ambient
 * medium) is taken in closed form: a packet crossing into it leaves back
 * or through in one boundary event, with the same probabilities and
 * directions as the bounce-by-bounce simulation; only the boundary-hit
 * count differs. */
#define ISMC_FLAG_SLIDE_BOUNCES 0x8u

/* One integrating sphere (or, when !present, the angle threshold that
 * separates specular/unscattered from diffuse light; CUDAISMLMC.h:413-433). */
typedef struct {
    int32_t present;
    int32_t _pad;
    double pw;     /* wall reflectance                                    */
    double fs;     /* sample port fraction                                */
    double fp;     /* source (R) / optional (T) port fraction             */
    double fd;     /* detector port fraction                              */
    double f;      /* total port fraction, normally fs + fp + fd          */
    double angle;  /* radians; used only when !present                    */
} ismc_sphere;

/* Sample between the spheres: layers in order from the illuminated face.
 * mu_s == 0 marks a non-scattering glass slide (CUDAISMLMC.h:102,155).
 * mu_t is taken as given, like the reference (normally mu_a + mu_s). */
typedef struct {
    int32_t n_layers;
    int32_t specular_included;   /* beam 8.1 deg off normal (CUDAISMLMC.h:347-352) */
    double n_ambient;            /* 1.0 in the reference                           */
    double roulette_threshold;   /* 1e-4 (CUDAISMLMC.h:140), 1e-3 in ISMC.py:136  */
    double roulette_m;           /* survivor weight factor, 10                     */
    double n[ISMC_MAX_LAYERS];
    double g[ISMC_MAX_LAYERS];
    double t[ISMC_MAX_LAYERS];
    double mu_a[ISMC_MAX_LAYERS];
    double mu_s[ISMC_MAX_LAYERS];
    double mu_t[ISMC_MAX_LAYERS];
    ismc_sphere R;
    ismc_sphere T;
} ismc_config;

/* Tallies, order A, R_d, R_s, T_d, T_u (CUDAISMLMC.h:754).  sum[k] is the
 * sum over packets of the weight fraction that packet left in tally k;
 * sumsq[k] the sum of squares (for the standard error).  Divide by
 * n_packets for the reference's fractions. */
typedef struct {
    double sum[5];
    double sumsq[5];
    uint64_t events[5];   /* interactions, boundary hits, entrances,
                             sphere decisions, re-injections            */
    uint64_t n_packets;
    uint64_t n_aborted;   /* packets stopped by the per-packet event cap
                             (2^22 events; not tallied).  Reached only by
                             genuinely long histories, e.g. a thin, weakly
                             scattering layer denser than both neighbours,
                             which guides packets by total reflection */
} ismc_tally;

/* Launch shape; zero fields mean "use the context's default".  Tallies do
   not depend on any of these fields. */
typedef struct {
    int32_t blocks_per_cu;
    int32_t chunk_packets;   /* packets per work-queue chunk */
    int32_t use_graph;       /* replay through a captured hipGraph when shapes repeat */
    int32_t idle_lanes;      /* lanes of a wave waiting for an entrance, sphere decision
                                or new packet before the wave serves them (1..64;
                                default 10, 12 when a point has a glass slide) */
    int32_t bound_lanes;     /* lanes of a wave waiting at an interface before the wave
                                serves them (1..64; default 12, 16 when a point has a
                                glass slide) */
    int32_t _reserved;
} ismc_tune;

typedef struct {
    char name[256];
    char arch[64];
    int32_t n_cu;
    int32_t clock_khz;
    int64_t total_mem;
    int32_t warp_size;
    int32_t _pad;
} ismc_device_info;

typedef struct ismc_ctx ismc_ctx;

int ismc_create(int device, ismc_ctx** out);
void ismc_destroy(ismc_ctx* ctx);
const char* ismc_last_error(const ismc_ctx* ctx);
int ismc_device_info_get(ismc_ctx* ctx, ismc_device_info* out);

/* Check a configuration without touching a GPU; the message (if msg is
 * non-NULL) explains a rejection.  Returns 0 or ISMC_EINVAL / ISMC_ERANGE. */
int ismc_validate_config(const ismc_config* cfg, char* msg, int32_t msg_len);

/* One configuration, packets first_packet_id .. first_packet_id+n-1 of
 * Philox stream (seed, stream).  Results do not depend on launch shape.
 * At most 2^31 packets per configuration per call (ISMC_ERANGE beyond:
 * the 64-bit fixed-point tallies, 2^-32 per unit of weight, keep twice
 * the headroom that many packets need); split larger runs into
 * packet-id ranges, whose tallies add exactly. */
int ismc_simulate(ismc_ctx* ctx, const ismc_config* cfg, uint64_t n_packets, uint64_t seed,
                  uint64_t first_packet_id, uint32_t stream, uint32_t flags, ismc_tally* out);

/* Many configurations with uneven packet counts in one persistent launch.
 * first_packet_id and streams may be NULL (0 and the point index). */
int ismc_simulate_batch(ismc_ctx* ctx, const ismc_config* cfgs, const uint64_t* n_packets,
                        const uint64_t* first_packet_id, const uint32_t* streams,
                        int32_t n_points, uint64_t seed, uint32_t flags, const ismc_tune* tune,
                        ismc_tally* out);

/* No counterpart in the reference: for callers that cannot allocate
 * (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  ismc_simulate_batch without any allocation: the
 * launch's device buffers (point tables, queue records, tally slots,
 * queue counter) live in a static device array of
 * ISMC_STATIC_SCRATCH_BYTES that every GPU the library runs on holds until
 * the process ends, and the host side stages through a pinned buffer that
 * ismc_create allocated.  At most ISMC_STATIC_MAX_POINTS configurations
 * (FP32 or FP64) per call, else ISMC_ERANGE before any device work; the
 * other limits and the tallies are those of ismc_simulate_batch, bit for
 * bit.  tune->use_graph is ignored (the call runs eagerly).  All calls on
 * one GPU share its array: calls that overlap on one GPU (from two
 * contexts or threads) race.  ISMC_EDEVICE if the array's address could
 * not be read at ismc_create. */
#define ISMC_STATIC_MAX_POINTS 1024
#define ISMC_STATIC_SCRATCH_BYTES 2982144
int ismc_simulate_batch_static(ismc_ctx* ctx, const ismc_config* cfgs, const uint64_t* n_packets,
                               const uint64_t* first_packet_id, const uint32_t* streams,
                               int32_t n_points, uint64_t seed, uint32_t flags,
                               const ismc_tune* tune, ismc_tally* out);

/* Per-packet records {A, R_d, R_s, T_d, T_u, events} (n_packets x 6
 * doubles) for one configuration: a debugging and parity aid. */
int ismc_simulate_packets(ismc_ctx* ctx, const ismc_config* cfg, uint64_t n_packets,
                          uint64_t seed, uint64_t first_packet_id, uint32_t stream,
                          uint32_t flags, double* per_packet, ismc_tally* out);

/* Device time (ms, hipEvents on the context's stream) of the last
 * transport launch, and the launch shape it used. */
int ismc_last_launch(ismc_ctx* ctx, float* kernel_ms, int32_t* grid_blocks,
                     int32_t* block_threads, int64_t* n_chunks);

/* How many launch graphs were captured and how many graph replays ran
 * (with ismc_tune.use_graph = 1 a batch call's launch sequence is captured
 * the second time its shape, buffers and seed come, and replayed from the
 * third on; the first runs without a graph). */
int ismc_graph_stats(ismc_ctx* ctx, int64_t* captures, int64_t* launches);

/* Evaluate the device physics helpers on host inputs (parity aid):
 * op 0 fresnel(a,b,c)->out[0], cos_t(out[1]); op 1 hg_cos(g=a,u=b);
 * op 2 cos(2 pi a); op 3 -ln(a); op 4 entrance reflectance of cfg at mu_z=a
 * (side R if a > 0 else T); op 5 spin: mu_z'(mu_z=a, g=b, u_hg=c, u_phi=d). */
int ismc_eval_physics(ismc_ctx* ctx, int32_t op, const ismc_config* cfg, const double* in4,
                      int32_t n, uint32_t flags, double* out2);

/* ---------------------------------------------------------------------
 * Surrogate MLP (no counterpart in the reference: it seeds the inverse
 * solver).  n_in -> [hidden SiLU] x depth -> n_out, BF16 MFMA with FP32
 * accumulation, FP32 master weights, Adam.  Parameters are exchanged in a
 * flat logical layout: for each layer W[in][out] row-major, then b[out].
 * Batches are multiples of 64 rows, at most max_batch.
 * --------------------------------------------------------------------- */
typedef struct ismc_mlp ismc_mlp;
int ismc_mlp_create(int device, int32_t n_in, int32_t hidden, int32_t depth, int32_t n_out,
                    int32_t max_batch, ismc_mlp** out);
void ismc_mlp_destroy(ismc_mlp* mlp);
const char* ismc_mlp_last_error(const ismc_mlp* mlp);
int64_t ismc_mlp_n_params(const ismc_mlp* mlp);
int ismc_mlp_set_params(ismc_mlp* mlp, const float* flat);
int ismc_mlp_get_params(ismc_mlp* mlp, float* flat);
/* y[n][n_out] = f(x[n][n_in]) (any n; padded internally) */
int ismc_mlp_forward(ismc_mlp* mlp, const float* x, int32_t n, float* y);
/* gradient of the mean squared error over the batch, flat logical layout */
int ismc_mlp_gradients(ismc_mlp* mlp, const float* x, const float* y, int32_t batch,
                       float* grads, float* loss);
/* keep a training set resident in device memory */
int ismc_mlp_load_dataset(ismc_mlp* mlp, const float* x, const float* y, int64_t n);
/* epochs x steps_per_epoch Adam steps on batches drawn in the order perm */
int ismc_mlp_train(ismc_mlp* mlp, const int32_t* perm, int32_t epochs, int32_t steps_per_epoch,
                   int32_t batch, float lr, float beta1, float beta2, float eps,
                   float* epoch_loss, float* total_ms);
/* one Adam step on a host batch (upload, forward, backward, update); *loss
 * = the batch's mean squared error before the step */
int ismc_mlp_train_step(ismc_mlp* mlp, const float* x, const float* y, int32_t batch, float lr,
                        float beta1, float beta2, float eps, float* loss);
int ismc_mlp_last_ms(ismc_mlp* mlp, float* ms);

/* ---------------------------------------------------------------------
 * Lookup-table uncertainty map.  Replaces the per-cell Monte Carlo loop
 * of makelookup (Lookup/lookup.py:347-376, with find_nearest,
 * lookup.py:296-299): for every cell (i, j) of P x P tables indexed
 * [i (R)][j (T)] whose Rstd and Tstd are finite and >= 0, draw M samples
 * R ~ N(Rvals[i], Rstd), T ~ N(Tvals[j], Tstd), look each up at the
 * nearest grid point, and return 100 * std / centre of the mu_a and mu_s'
 * values found (NaN elsewhere).  Rvals and Tvals must be the same uniform
 * increasing grid.  All pointers are host memory; the call blocks.
 * Normals from Philox4x32-10 with key = seed (Box-Muller, FP64).
 * --------------------------------------------------------------------- */
int ismc_table_uncertainty(int device, const double* mua, const double* mus, const double* rstd,
                           const double* tstd, const double* Rvals, const double* Tvals, int32_t P,
                           int32_t M, uint64_t seed, double* mua_unc, double* mus_unc);
const char* ismc_table_last_error(void);

/* For each of M measurements (R[m], T[m]) the index of the cell of a
 * G-cell grid of predicted readings (pR, pT) with the least relative
 * misfit ((pR - R)/R)^2 + ((pT - T)/max(T, 1e-6))^2: numpy's argmin over
 * the grid, first index among equal minima, a NaN cost first (no
 * counterpart in the reference: the global search that seeds the inverse
 * solver from the surrogate).  FP64, host pointers, blocks; errors in
 * ismc_table_last_error. */
int ismc_grid_argmin(int device, const float* pR, const float* pT, int32_t G, const double* R,
                     const double* T, int32_t M, int32_t* idx);

#ifdef __cplusplus
}
#endif
#endif
