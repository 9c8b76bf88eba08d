// libismc — photon transport for MI355X (gfx950).
//
// One persistent kernel runs every packet history of a launch: a packet
// enters the sample from one of the two integrating spheres, hops, drops
// and spins through the layers (Henyey-Greenstein scattering, Fresnel
// reflection/refraction at every interface, Russian roulette), leaves
// into a sphere, and is either captured (specular/unscattered port or
// diffuse wall) or returned to the sample, possibly from the other side.
// Semantics follow CUDA/CUDAISMLMC.h:99-654 and ISMLMC.py:101-347 +
// MLMC.py:76-202 of the reference; the structure is our own:
//
//  * Event machine.  Each loop iteration advances every lane in flight by
//    several events of its packet, one per flight step: a flight that ends
//    in an interaction (hop, drop, roulette, spin) or on an interface.  A
//    lane that reaches an interface records its event's draws and waits;
//    the interface code (Fresnel reflect / refract / exit, layer change)
//    runs for all waiting lanes at once, before a flight step, when enough
//    of them wait (the bound-lane threshold).
//    The rare steps (an entrance through the outer face, a sphere
//    decision, the end-of-packet tally, taking a new packet) likewise wait
//    for the idle-lane threshold and run together in one "rare pass".
//    With 64 lanes nearly every iteration has some lane at an interface and
//    some lane at a rare step, so running those paths for many lanes at
//    once instead of on every iteration removes most of the divergent
//    work.  A wave never waits for its longest history: finished lanes
//    take new packets at the next rare pass.
//  * Only z and mu_z are carried.  The slab is laterally infinite and
//    every tally, boundary distance, Fresnel test and port test depends on
//    z and mu_z alone; after a spin the new mu_z depends only on (mu_z,
//    theta, cos phi).  x, y, mu_x, mu_y (and sin phi) are dead state in the
//    reference and are not computed.
//  * Random numbers: Philox4x32-10 (Salmon et al. SC'11), key = seed,
//    counter = (packet id lo, event index, packet id hi, stream).  The event
//    index sits in the counter word that the first round does not
//    multiply, so four of a block's twenty 32 x 32 -> 64 products depend on
//    the packet alone and are taken once, when a lane takes the packet
//    (philox_fixed); each event's block then costs sixteen.  Each
//    event draws one 4-word block.  Words are used as
//      entrance:     w0 reflect test, w1 first step (or return draw), w2 new |mu_z|
//      interaction:  w0 next step, w1 HG, w2 azimuth, w3 roulette
//      interface:    w0 Fresnel test, w1 return draw, w2 new |mu_z| (on exit)
//    so a history is a pure function of (seed, stream, packet id):
//    results do not depend on launch shape or scheduling.
//  * Work queue.  Every configuration ("point") is split into chunks of
//    packet ids, numbered point after point.  Each wave takes chunk numbers
//    from a global atomic counter and finds the chunk's point by a binary
//    search over the points' first chunk numbers (one 32-byte range record
//    per point, so the host writes and uploads O(points), not O(chunks));
//    the chunk's point table (layers + sphere constants) is copied into a
//    per-wave LDS slot (two slots, so lanes can start the next chunk while
//    others finish the last one).  The current layer's constants live in
//    registers and are reloaded from LDS only when a packet crosses into
//    another layer.
//  * Tallies.  Per-packet sums are kept in the transport precision; at the
//    end of a packet its A and exit weight (and their squares) are added as
//    64-bit fixed point (2^-32 quantum, round to nearest) into a per-wave
//    LDS accumulator of its chunk's slot; a drained slot is flushed with
//    one 64-bit atomic per tally into the point's global slot.  Integer
//    sums are exact and order independent, so tallies are bitwise
//    reproducible; the host converts them to FP64.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/ismc.h"
#include "philox.hpp"

namespace ismc {

constexpr int kMaxLayers = ISMC_MAX_LAYERS;
constexpr int kBlock = 256;
constexpr int kWavesPerBlock = kBlock / 64;
constexpr uint32_t kMaxEventsPerPacket = 1u << 22;
// Packets of one point in one call.  A packet adds at most about 1 (its
// weight, a little more after a roulette survival) to a tally in 2^-32
// fixed point, so a u64 slot holds 2^31 packets with twice that headroom;
// 2^32 packets of weight 1 would wrap it to 0.
constexpr uint64_t kMaxPacketsPerPoint = 1ull << 31;
// Flight steps per loop iteration.  More steps share one iteration's rare
// pass decision and ballots between more events; the interface pass runs
// before every step, so a lane that stopped at an interface in one step
// flies again in the next once enough lanes wait.  Measured on MI355X
// (slab between slides / three layers / one layer mu_s 10, FP32): 4 steps
// +4.5 % / +3.8 % / +12 % over 2, 6 steps a further +0.7 % / +0.8 % /
// -0.4 %, 8 steps no better on the slab and -7 % on mu_s 10; FP64 with 3
// steps +4 % over 1, with 4 a further +0.1 % / +0.3 % / +3.4 % (6: no gain).
template <typename R> struct FlightSteps { static constexpr int value = 4; };
template <> struct FlightSteps<float> { static constexpr int value = 6; };
// waves per SIMD the FP32 tally kernel is compiled for at least (it takes
// 67 VGPRs, no spills, so 7 fit; a bound of 8 caps it at 64 VGPRs with 12
// bytes per lane spilled); the FP64
// kernel needs about twice the registers and the per-packet output
// variant a few more, both are left to the compiler
template <typename R, bool kPerPacket> struct MinWaves { static constexpr int value = 1; };
template <> struct MinWaves<float, false> { static constexpr int value = 7; };
constexpr int kSlotWords = 16;  // per point: sum[5], sumsq[5], events[5], aborted
// A wave's LDS tally accumulator is kept in kAccCopies interleaved copies
// (word k of copy c at k * kAccCopies + c, so a word's copies sit in
// consecutive 8-byte slots): a lane that ends a packet adds to copy
// lane % kAccCopies.  The lanes ending in one rare pass all add to the same
// words, and same-address LDS atomics of one instruction serialise; four
// copies spread them over four addresses.  Measured on one layer with
// mu_s 10 (short histories, many packet ends per step): 17.6 % less kernel
// time with four copies, the same with eight; the longer workloads within
// 0.2 % with four and up to 1.4 % slower with eight (more registers).  A
// slot's flush sums the copies.
constexpr int kAccCopies = 4;
constexpr int kAccWords = kSlotWords * kAccCopies;

// Layer kinds.  A slide is a glass layer whose neighbours are scattering
// layers or the ambient medium: the kernel takes a packet's bounces inside
// it in closed form (the interface pass and the entrance below); any other
// glass layer is flown face by face.  Glass tests read kind >= KIND_GLASS.
enum : int32_t { KIND_HG = 0, KIND_ISO = 1, KIND_GLASS = 2, KIND_SLIDE = 3 };

// A packet that refracts into a glass slide from a less dense medium a at
// grazing incidence lands on the critical angle of the glass->a face.  In
// finite precision 1 - mu^2 can round so that the way back reads as total
// internal reflection; with TIR at the far face too (glass->air) the packet
// would bounce in the non-absorbing slide for ever (the FP32 reference path
// has no exit from that loop).  In a run of adjacent glass layers the
// rounding also accumulates: n^2 (1 - mu^2) is kept only to the last bit at
// each glass-glass crossing, and where the two directions' index ratios
// round apart it drifts the same way at every round trip, until neither
// face that bounds the run lets the packet out.  The refracted direction in
// a glass layer is therefore kept at least GlassGuard (relative, in sin^2)
// inside the critical cone of the layer's exit index (glass_guard_index):
// on one side of it every index out to the medium that bounds its run of
// glass layers (a scattering layer or the ambient medium) is at least that
// large, so the packet can leave the run there.  For a slide entered from a
// scattering layer at least as dense as the medium beyond the slide this is
// the cone of the face the packet came through; it moves only packets
// within ~sqrt(GlassGuard) of grazing incidence.
template <typename R> struct GlassGuard;
template <> struct GlassGuard<float> { static constexpr float value = 9.5367431640625e-07f; };  // 2^-20
template <> struct GlassGuard<double> { static constexpr double value = 9.094947017729282e-13; };  // 2^-40
// Phases of a lane.  The ones that wait for a rare pass (need a packet,
// enter, sphere decision, end) are exactly those with bit 1 clear, so one
// AND and one compare find them.
enum : int32_t { PH_NEED = 0, PH_ENTER = 1, PH_FLIGHT = 2, PH_DONE = 3, PH_SPHERE = 4, PH_END = 5,
                 PH_BOUND = 6 };
static_assert((PH_NEED & 2) == 0 && (PH_ENTER & 2) == 0 && (PH_SPHERE & 2) == 0 && (PH_END & 2) == 0 &&
                  (PH_FLIGHT & 2) && (PH_DONE & 2) && (PH_BOUND & 2),
              "rare-pass phases are the ones with bit 1 clear");

// Constants of the interface a packet meets when it leaves a layer in one
// direction (towards the neighbouring layer, or the ambient medium at the
// stack's face), derived on the host in FP64: they depend on the two
// indices only, so no hit recomputes them.
template <typename R>
struct alignas(16) Iface {
  R eta;   // n_here / n_there (Snell)
  R qq;    // ((n_there - n_here) / (n_there + n_here))^2, normal incidence
  R cmin;  // entering a glass layer denser than its exit index
           // (glass_guard_index): the least |mu_z| inside it
           // (glass_entry_cos, GlassGuard); 0 otherwise
  R mut;   // rescales the remaining step s on transmission: mu_t here /
           // mu_t there when both layers scatter, 0 out of a glass layer, 1
           // into one; -1 where the packet leaves the stack (the interface
           // pass reads "outside" from its sign, no layer count); +inf into
           // a slide, whose bounces the pass takes in closed form (there
           // the step is set, not rescaled)
};

// Constants of one layer, derived on the host in FP64.  n and mu_t are not
// read on the device (they enter through out[]); the layout without them
// measured 0.5 % slower on MI355X (LDS offsets of the layer reads), so
// they stay as the layer's record.
template <typename R>
struct alignas(16) Layer {
  R half, zc, n, step_k;     // half thickness; centre depth (record only); step_k = ln2 / mu_t:
                             // s = -log2(u) * step_k
  R absfrac, hg_a, hg_b, hg_c;  // mu_a/mu_t; scattering constants (Draw<R>::constants)
  R hg_d, hg_e, mu_t;
  int32_t kind;
  Iface<R> out[2];           // leaving towards lower z (0) / higher z (1)
};

// Everything a packet needs about its configuration.
template <typename R>
struct alignas(16) Point {
  int32_t n_layers;
  int32_t first_layer[2];    // entered from the R side / T side
  int32_t ent_glass[2];      // first layer on that side is a glass slide
  int32_t cuda_roulette;
  R n_amb, t_tot, roul_thr, roul_m;  // n_amb: record only (see Layer)
  R roul_p, mu0_z;
  R cos_spec[2];             // |mu_z| >= cos_spec: specular (R) / unscattered (T)
  R p_return[2];             // sphere return probability; -1 without a sphere
  R ent_eta1[2], ent_qq1[2]; // entrance from the ambient medium into the first layer
  R ent_eta2[2], ent_qq2[2]; // and from a first-layer slide into the layer behind it
  R ent_cmin[2];             // glass_entry_cos bound of that entrance (0: none)
  Layer<R> layer[kMaxLayers];
};

// a point's packets in the launch's queue
struct alignas(16) Range {
  uint64_t first_pid;  // packet id of the point's first packet
  uint64_t out_base;   // dense index of that packet (per-packet output)
  uint64_t n_packets;
  uint32_t chunk0;     // number of the point's first chunk
  uint32_t stream;
};

struct Args {
  const void* points;        // Point<R>[n_points]
  const Range* ranges;       // [n_points]
  uint32_t n_points;
  uint32_t n_chunks;
  uint32_t chunk_packets;
  uint32_t seed_lo, seed_hi;
  uint32_t* chunk_counter;
  unsigned long long* slots; // [n_points][kSlotWords]
  double* per_packet;        // optional [n][6]
  uint32_t idle_lanes;       // lanes waiting for a rare step that trigger a rare pass
  uint32_t bound_lanes;      // lanes waiting at an interface that trigger an interface pass
};

// ---------------------------------------------------------------------------
// device math
template <typename R> struct DM;
template <> struct DM<float> {
  static __device__ __forceinline__ float rcp(float x) { return __builtin_amdgcn_rcpf(x); }
  static __device__ __forceinline__ float sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
  static __device__ __forceinline__ float nlog2(float u) { return -__builtin_amdgcn_logf(u); }
  static __device__ __forceinline__ float cos2pi(float u) { return __builtin_amdgcn_cosf(u); }
  static __device__ __forceinline__ float abs(float x) { return __builtin_fabsf(x); }
  static __device__ __forceinline__ float max0(float x) { return __builtin_fmaxf(x, 0.0f); }
  static __device__ __forceinline__ float csign(float a, float b) { return __builtin_copysignf(a, b); }
  // round(v 2^32) for 0 <= v < 2^32 as the integer part (a truncating
  // convert) and round(fract(v) 2^32), which is below 2^32 - 255 for an FP32
  // fraction: the same value as rint(v 2^32) converted to 64 bits (adding
  // the even number floor(v) 2^32 changes neither the rounding nor its tie
  // rule), in five instructions instead of seven
  static __device__ __forceinline__ unsigned long long fx(float v) {
    const uint32_t hi = (uint32_t)v;
    const uint32_t lo = (uint32_t)__builtin_rintf(__builtin_amdgcn_fractf(v) * 4294967296.0f);
    return ((unsigned long long)hi << 32) | lo;
  }
};
template <> struct DM<double> {
  static __device__ __forceinline__ double rcp(double x) { return 1.0 / x; }
  static __device__ __forceinline__ double sqrt(double x) { return __builtin_sqrt(x); }
  static __device__ __forceinline__ double nlog2(double u) { return -::log2(u); }
  static __device__ __forceinline__ double cos2pi(double u) { return ::cospi(2.0 * u); }
  static __device__ __forceinline__ double abs(double x) { return __builtin_fabs(x); }
  static __device__ __forceinline__ double max0(double x) { return __builtin_fmax(x, 0.0); }
  static __device__ __forceinline__ double csign(double a, double b) { return __builtin_copysign(a, b); }
  static __device__ __forceinline__ unsigned long long fx(double v) {
    return (unsigned long long)__builtin_rint(v * 4294967296.0);
  }
};

using ismc_rng::PhiloxFixed;
using ismc_rng::philox_block;
using ismc_rng::philox_fixed;

// 23-bit uniform 1 - (x >> 9) 2^-23 in (0, 1], exact in float and double.
// One v_alignbit_b32 shifts the word's top 23 bits under the exponent of
// 1.0 (the float 1 + (x >> 9) 2^-23 in [1, 2)), and 2 - that is exact
// (Sterbenz): two full-rate instructions where shift, convert and fma took
// three; measured 1.6 % faster on the slab between slides.
template <typename R>
__device__ __forceinline__ R u01(uint32_t x) {
  const uint32_t m = __builtin_amdgcn_alignbit(0x7Fu, x, 9u);  // 0x3f800000 | (x >> 9)
  return (R)(2.0f - __builtin_bit_cast(float, m));
}

// A draw as the flight step keeps it.  FP32 keeps the float m = 2 - u =
// 1 + (x >> 9) 2^-23 in [1, 2) (one v_alignbit_b32) and forms u = 2 - m
// only where u itself is read: the Henyey-Greenstein and azimuth formulas
// take m directly, 1 - g + 2g u = (1 + 3g) - 2g m, 2u - 1 = 3 - 2m, and
// cos 2 pi u = cos 2 pi m, which saves the subtraction of two of each
// event's three uniforms.  FP64 keeps u and the formulas written with it,
// operation for operation as the oracle has them.
//
// The FP32 scattering cosine is one formula for both kinds of layer,
// cos(theta) = H - K r^2 + F m with r = 1 / ((1 + 3g) - 2g m):
//   g != 0: H = (1 + g^2) / (2g), K = (1 - g^2)^2 / (2g), F = 0, which is
//           the Henyey-Greenstein inverse CDF (1 + g^2 - q^2) / (2g) with
//           q = (1 - g^2) r (CUDAISMLMC.h:284-285);
//   g == 0: H = 3, K = 0, F = -2 (and r = 1), i.e. 3 - 2m = 2u - 1, the
//           isotropic cosine (:284).
// No per-event compare and select of the layer's kind; the layer record
// holds (hg_a .. hg_e) = (H, K, 1 + 3g, 2g, F).  FP64 holds (1 + g^2,
// 1 - g^2, 1 - g, 2g, 1/(2g)) and branches on the kind as the oracle does.
template <typename R> struct Draw;
template <> struct Draw<float> {
  static __device__ __forceinline__ float keep(uint32_t x) {
    return __builtin_bit_cast(float, __builtin_amdgcn_alignbit(0x7Fu, x, 9u));
  }
  static __device__ __forceinline__ float u(float m) { return 2.0f - m; }
  static __host__ __device__ void constants(double g, float* c) {
    if (g == 0.0) {
      c[0] = 3.0f; c[1] = 0.0f; c[2] = 1.0f; c[3] = 0.0f; c[4] = -2.0f;
      return;
    }
    const double h = 1.0 / (2.0 * g), k = 1.0 - g * g;
    c[0] = (float)((1.0 + g * g) * h);
    c[1] = (float)(k * k * h);
    c[2] = (float)(1.0 + 3.0 * g);
    c[3] = (float)(2.0 * g);
    c[4] = 0.0f;
  }
  static __device__ __forceinline__ float hg_cos(int, float hh, float hk, float hc3, float hd, float hf,
                                                 float m) {
    const float r = DM<float>::rcp(__builtin_fmaf(-hd, m, hc3));
    return __builtin_fmaf(hf, m, __builtin_fmaf(-hk, r * r, hh));
  }
  // the kept form of a uniform u given as a double (eval_kernel)
  static __device__ __forceinline__ float from_u(double u) { return (float)(2.0 - u); }
  static __device__ __forceinline__ float cos2pi(float m) { return DM<float>::cos2pi(m); }
};

// Unpolarised Fresnel reflectance for light going from index ni to nt at
// direction cosine mu_z (fresnel_snell, CUDAISMLMC.h:79-96), written with
// cosines instead of angles: rs = (ni ci - nt ct)/(ni ci + nt ct),
// rp = (ni ct - nt ci)/(ni ct + nt ci), R = (rs^2 + rp^2)/2.  Same
// near-normal shortcut (|mu_z| > 0.99999) and total-internal-reflection rule.
// ct (the transmitted cosine, Snell) is returned for the refraction.
// Both amplitudes are divided through by nt, so only eta = ni/nt and the
// normal-incidence value qq = ((nt - ni)/(nt + ni))^2 enter; the host
// derives both per interface.
template <typename R>
__device__ __forceinline__ R fresnel(R eta, R qq, R mu_z, R& ct) {
  const R ci = DM<R>::abs(mu_z);
  // ct^2 = 1 - sin^2(theta_t); not positive: total internal reflection
  const R ct2 = (R)1 - eta * eta * ((R)1 - ci * ci);
  ct = DM<R>::sqrt(fmax((R)0, ct2));
  // rs = (a - ct) / (a + ct), rp = (c - ci) / (c + ci) and R = (rs^2 +
  // rp^2) / 2 over one reciprocal: a transcendental takes four full-rate
  // slots, and B D >= min(ci, ct)^2 stays far above FP32's underflow
  const R a = eta * ci, c = eta * ct;
  const R A = a - ct, B = a + ct, C = c - ci, D = c + ci;
  const R AD = A * D, CB = C * B, BD = B * D;
  R Rf = (AD * AD + CB * CB) * DM<R>::rcp((R)2 * (BD * BD));
  Rf = ct2 <= (R)0 ? (R)1 : Rf;
  Rf = ci > (R)0.99999 ? qq : Rf;
  return Rf;
}

// Entrance reflectance (incident_reflection, CUDAISMLMC.h:99-122): through a
// glass slide r1 + (1-r1)^2 r2 / (1 - r1 r2) with both Fresnel terms at the
// air-side cosine; otherwise one air/sample interface.
template <typename R>
__device__ __forceinline__ R entrance_R(const Point<R>& P, int side, R mu_z, R& ct1) {
  R r = fresnel<R>(P.ent_eta1[side], P.ent_qq1[side], mu_z, ct1);
  if (P.ent_glass[side]) {  // the slide's far face only where there is a slide
    R ct2;
    const R r2 = fresnel<R>(P.ent_eta2[side], P.ent_qq2[side], mu_z, ct2);
    // r = 1 (an ambient medium denser than the slide, at grazing incidence)
    // reflects totally; with r2 = 1 too the series would read 0/0
    r = r < (R)1 ? r + ((R)1 - r) * ((R)1 - r) * r2 * DM<R>::rcp((R)1 - r * r2) : (R)1;
  }
  return r;
}

// Henyey-Greenstein cos(theta) (CUDAISMLMC.h:284-285).
template <typename R>
__device__ __forceinline__ R hg_cos(int kind, R ha, R hb, R hc, R hd, R he, R u) {
  const R q = hb * DM<R>::rcp(hc + hd * u);
  const R c = (ha - q * q) * he;
  return kind == KIND_ISO ? (R)2 * u - (R)1 : c;
}

template <> struct Draw<double> {
  static __device__ __forceinline__ double keep(uint32_t x) { return u01<double>(x); }
  static __device__ __forceinline__ double u(double v) { return v; }
  static __host__ __device__ void constants(double g, double* c) {
    c[0] = 1 + g * g;
    c[1] = 1 - g * g;
    c[2] = 1 - g;
    c[3] = 2 * g;
    c[4] = g == 0.0 ? 0.0 : 1 / (2 * g);
  }
  static __device__ __forceinline__ double hg_cos(int kind, double ha, double hb, double hc, double hd,
                                                  double he, double u) {
    return ::ismc::hg_cos<double>(kind, ha, hb, hc, hd, he, u);
  }
  static __device__ __forceinline__ double from_u(double u) { return u; }
  static __device__ __forceinline__ double cos2pi(double u) { return DM<double>::cos2pi(u); }
};

// New mu_z after a scattering by theta with azimuth 2 pi u_phi, the
// azimuth given as the draw the flight step keeps (CUDAISMLMC.h:287-308,
// z component).  The reference takes mu_z' = sign(mu_z) cos_t where
// |mu_z| > 0.99999, because its x and y components divide by
// sqrt(1 - mu_z^2).  The z component alone has no such division: the
// general form is exact there too (at |mu_z| = 1 it is that same
// sign(mu_z) cos_t) and differs from the shortcut by at most
// sqrt(2e-5) sin(theta) |cos(phi)|, zero on average over phi.  The FP32
// kernel uses the general form everywhere (three instructions fewer per
// interaction, 2.5 % faster; slab tallies unchanged to six digits at
// 2^29 packets); the FP64 kernel keeps the shortcut, operation for
// operation as the oracle has it.
template <typename R>
__device__ __forceinline__ R spin_z(R mu_z, R cos_t, R u_phi) {
  const R cphi = Draw<R>::cos2pi(u_phi);
  const R st_sz = DM<R>::sqrt(fmax((R)0, ((R)1 - cos_t * cos_t) * ((R)1 - mu_z * mu_z)));
  const R general = mu_z * cos_t - st_sz * cphi;
  if constexpr (sizeof(R) == 4) return general;
  // near-vertical: mu_z' = sign(mu_z) * cos_t (cos_t keeps its own sign)
  return DM<R>::abs(mu_z) > (R)0.99999 ? (mu_z > (R)0 ? cos_t : -cos_t) : general;
}

// cosine inside a glass layer of index ng, clamped away from the critical
// angle of its exit index nx < ng (glass_guard_index, GlassGuard): at least
// cmin = sqrt(1 - (1 - guard) (nx/ng)^2),
// which the host derives per interface (Iface::cmin, Point::ent_cmin; 0
// where no bound applies)
template <typename R>
__device__ __forceinline__ R glass_entry_cos(R ct, R cmin) {
  return ct > cmin ? ct : cmin;  // a compare and a select (fmax adds two canonicalizations)
}

// LDS records of a lane's point and layer.  Their byte offsets are 32-bit
// and the indices small, so 24-bit multiplies (full rate) form them; plain
// indexing compiles to v_mad_u64_u32 / v_mul_lo_u32 (half / quarter rate).
template <typename R>
__device__ __forceinline__ const Point<R>& point_at(const Point<R>* pts, int set) {
  return *reinterpret_cast<const Point<R>*>(reinterpret_cast<const char*>(pts) +
                                            __umul24((uint32_t)set, (uint32_t)sizeof(Point<R>)));
}
// A lane keeps its current layer as the byte offset of that layer's record
// from the wave's point slots (slot set, layer index): an interface reads
// its record and steps to the neighbouring layer by adding offsets.
template <typename R>
__device__ __forceinline__ uint32_t layer_offset(int set, int layer) {
  return __umul24((uint32_t)set, (uint32_t)sizeof(Point<R>)) + (uint32_t)offsetof(Point<R>, layer) +
         __umul24((uint32_t)layer, (uint32_t)sizeof(Layer<R>));
}
template <typename R>
__device__ __forceinline__ const Layer<R>& layer_rec(const Point<R>* pts, uint32_t lofs) {
  return *reinterpret_cast<const Layer<R>*>(reinterpret_cast<const char*>(pts) + lofs);
}

__device__ __forceinline__ uint32_t lane_rank(uint64_t mask) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                   __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

template <typename R>
__device__ __forceinline__ void copy_point(Point<R>* dst, const Point<R>* src, int lane) {
  const uint32_t* s = reinterpret_cast<const uint32_t*>(src);
  uint32_t* d = reinterpret_cast<uint32_t*>(dst);
  constexpr int words = (int)(sizeof(Point<R>) / 4);
  for (int i = lane; i < words; i += 64) d[i] = s[i];
}

template <typename R>
__device__ __forceinline__ void flush_slot(unsigned long long* acc, unsigned long long* slots,
                                           uint32_t point, int lane) {
  // The point's row address is wave-uniform (scalar registers).  The lane
  // index passes through an empty asm operand so that the compiler cannot
  // hoist a 64-bit per-lane address out of the kernel's loop (it would
  // hold two VGPRs for the whole kernel, or spill them); the asm emits no
  // instruction.
  unsigned long long* row = slots + (size_t)__builtin_amdgcn_readfirstlane(point) * kSlotWords;
  uint32_t l = (uint32_t)lane;
  asm volatile("" : "+v"(l));
  if (l < (uint32_t)kSlotWords) {
    // one copy at a time (once per chunk; an unrolled sum would hold every
    // copy in registers at the kernel's busiest point)
    unsigned long long v = 0ull;
#pragma unroll 1
    for (int c = 0; c < kAccCopies; ++c) v += acc[l * kAccCopies + c];
    if (v) atomicAdd(&row[l], v);
  }
  // every lane's reads are issued before any lane clears a copy (a wave's
  // LDS operations execute in order)
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  for (uint32_t i = l; i < (uint32_t)kAccWords; i += 64) acc[i] = 0ull;
}

// kSlides: the launch has a slide in some point.  The closed form of a
// slide's bounces is compiled into that instantiation only, so a launch
// without slides runs code without it (its extra branches cost 1-2 % on
// stacks without slides).
template <typename R, bool kPerPacket, bool kSlides>
__global__ __launch_bounds__(kBlock, (MinWaves<R, kPerPacket>::value)) void transport_kernel(Args a) {
  __shared__ Point<R> s_pt[kWavesPerBlock][2];
  __shared__ unsigned long long s_acc[kWavesPerBlock][2][kAccWords];

  const int wid = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  Point<R>* pts = s_pt[wid];
  unsigned long long* acc0 = s_acc[wid][0];
  unsigned long long* acc1 = s_acc[wid][1];
  for (int i = lane; i < kAccWords; i += 64) { acc0[i] = 0ull; acc1[i] = 0ull; }
  __builtin_amdgcn_wave_barrier();

  const Point<R>* gpts = reinterpret_cast<const Point<R>*>(a.points);
  const uint32_t k0 = a.seed_lo, k1 = a.seed_hi;
  const uint32_t idle_thr = a.idle_lanes;
  const uint32_t bound_thr = a.bound_lanes;

  // wave-uniform queue state
  int cur_set = 1;
  uint32_t set_point0 = 0xffffffffu, set_point1 = 0xffffffffu;
  uint32_t outstanding0 = 0, outstanding1 = 0;
  uint64_t win_next = 0, win_end = 0, win_first = 0, win_outbase = 0;
  uint32_t win_stream = 0;
  bool exhausted = false;

  // lane state
  int phase = PH_NEED, side = 0, set = 0, term = 0;
  uint32_t lofs = 0;  // the current layer (layer_offset)
  bool aborted = false;
  // A is the packet's starting weight plus the weight Russian roulette has
  // added or removed: the weight absorbed so far is A - w (every other
  // change of w is an absorption), so an interaction updates w alone
  // zr is the packet's depth relative to the centre of its layer, whose
  // faces are then at -half and +half exactly: a flight reaches the face
  // ahead when |zr + s mu_z| > half (two instructions), and a packet on a
  // face (zr = +-half) with a zero step stays inside
  R w = 0, zr = 0, muz = 1, s = 0, A = 0;
  R su1 = 0, su2 = 0;  // return draw and new |mu_z| of a pending sphere decision
  R bu0 = 0;           // Fresnel draw of a pending interface hit
  uint64_t pid = 0, outidx = 0;
  uint32_t ev = 0, stream = 0;
  PhiloxFixed pf{0u, 0u, 0u, 0u, 0u};  // the packet's event-independent Philox products
  // event counts of the packet; its interactions are ev - e_bnd - e_inc
  // (every event is an interaction, an interface hit or an entrance)
  uint32_t e_bnd = 0, e_inc = 0, e_sph = 0, e_rei = 0;
  // current layer cache
  R L_half = 0, L_step = 0, L_abs = 0, L_ha = 0, L_hb = 0, L_hc = 0, L_hd = 0,
    L_he = 0;
  int L_kind = KIND_GLASS;
  R L_thr = 0;  // the point's roulette threshold

  for (;;) {
    // Lanes in flight run their next event every flight step.  A flight
    // that reaches an interface only records its event's draws and waits;
    // the interface code (Fresnel, refraction, layer change) runs in an
    // "interface pass", before a flight step, once bound_thr lanes wait (or
    // none is in flight), after which those lanes go on with their next
    // flight event in that step.  The rare steps (sphere decision,
    // end-of-packet tally, refill, entrance) likewise run in a "rare pass"
    // at the top of an iteration once idle_thr lanes wait for one.  So the
    // wave does not pay for the code of every path on every step, and every
    // packet still draws the same Philox block for the same event.
    //
    // A packet in flight after more than kMaxEventsPerPacket events is
    // dropped (counted, not tallied) before it runs another.  An event
    // index grows by at most FlightSteps per iteration and the passes run
    // no event of their own, so only a wave with a lane within FlightSteps
    // of the limit checks, here, and runs a single flight event in this
    // iteration; every other wave does no per-event check.
    const bool near_limit =
        __ballot(ev + (uint32_t)FlightSteps<R>::value > kMaxEventsPerPacket) != 0;
    if (near_limit) {
      asm volatile("");  // keeps this rare block a branch, not selects run by every wave
      if (phase == PH_FLIGHT && ev > kMaxEventsPerPacket) {
        phase = PH_END;
        aborted = true;
      }
    }
    const uint64_t fly = __ballot(phase == PH_FLIGHT);
    const uint64_t bnd = __ballot(phase == PH_BOUND);
    const uint64_t idle = __ballot((phase & 2) == 0);  // waiting for a rare pass
    if ((fly | bnd | idle) == 0) break;
    const bool rare = idle != 0 && ((fly | bnd) == 0 || (uint32_t)__popcll(idle) >= idle_thr);
    const bool bpass0 = bnd != 0 && (fly == 0 || (uint32_t)__popcll(bnd) >= bound_thr);

    if (rare) {
      // ---- sphere capture / return (CUDAISMLMC.h:377-410, 442-475, 518-550, 583-615)
      if (phase == PH_SPHERE) {
        const Point<R>& P = point_at<R>(pts, set);
        e_sph += 1;
        if (DM<R>::abs(muz) >= P.cos_spec[side]) {
          term = side ? 4 : 2;
          phase = PH_END;
        } else if (Draw<R>::u(su1) < P.p_return[side]) {
          muz = side ? -Draw<R>::u(su2) : Draw<R>::u(su2);
          phase = PH_ENTER;
          e_rei += 1;
          if (ev > kMaxEventsPerPacket) { phase = PH_END; aborted = true; }
        } else {
          term = side ? 3 : 1;
          phase = PH_END;
        }
      }
      // ---- end of packet: tallies into the chunk slot's LDS accumulator
      const bool ending = (phase == PH_END);
      if (ending) {
        // this lane's copy of the slot's accumulator: word k at acc[k * kAccCopies]
        // (the copy index passes through an empty asm operand so that the
        // compiler forms the address here instead of holding two per-lane
        // addresses in registers through the whole kernel)
        uint32_t cpy = (uint32_t)lane & (uint32_t)(kAccCopies - 1);
        asm volatile("" : "+v"(cpy));
        unsigned long long* acc = (set ? acc1 : acc0) + cpy;
        A = aborted ? (R)0 : A - w;  // the absorbed weight
        constexpr int C = kAccCopies;
        if (aborted) {
          atomicAdd(&acc[15 * C], 1ull);
          term = 0;
        } else {
          atomicAdd(&acc[0 * C], DM<R>::fx(A));
          atomicAdd(&acc[5 * C], DM<R>::fx(A * A));
          if (term) {
            atomicAdd(&acc[term * C], DM<R>::fx(w));
            atomicAdd(&acc[(5 + term) * C], DM<R>::fx(w * w));
          }
        }
        atomicAdd(&acc[10 * C], (unsigned long long)(ev - e_bnd - e_inc));
        atomicAdd(&acc[11 * C], (unsigned long long)e_bnd);
        atomicAdd(&acc[12 * C], (unsigned long long)e_inc);
        atomicAdd(&acc[13 * C], (unsigned long long)e_sph);
        atomicAdd(&acc[14 * C], (unsigned long long)e_rei);
        if (kPerPacket) {
          double* o = a.per_packet + outidx * 6;
          o[0] = (double)A;
          o[1] = term == 1 ? (double)w : 0.0;
          o[2] = term == 2 ? (double)w : 0.0;
          o[3] = term == 3 ? (double)w : 0.0;
          o[4] = term == 4 ? (double)w : 0.0;
          o[5] = (double)ev;
        }
        e_bnd = e_inc = e_sph = e_rei = 0;
        term = 0;
        aborted = false;
        phase = PH_NEED;
      }
      const uint64_t end0 = __ballot(ending && set == 0);
      const uint64_t end1 = __ballot(ending && set == 1);
      outstanding0 -= (uint32_t)__popcll(end0);
      outstanding1 -= (uint32_t)__popcll(end1);

      // ---- refill idle lanes from the wave's window
      uint64_t need = __ballot(phase == PH_NEED);
      // The loop only hands out packet ids (a chunk switch can take it round
      // twice); the lanes it served are set up once after it, so that their
      // state is not carried round the loop (register copies at every pass)
      uint64_t fresh = 0;  // lanes given a packet in this pass
      while (need) {
        if (win_next == win_end) {
          if (exhausted) {
            if (phase == PH_NEED) phase = PH_DONE;
            break;
          }
          const int ns = cur_set ^ 1;
          const uint32_t busy = ns ? outstanding1 : outstanding0;
          if (busy) break;  // lanes of the chunk held in that slot still run
          const uint32_t sp = ns ? set_point1 : set_point0;
          if (sp != 0xffffffffu) flush_slot<R>(ns ? acc1 : acc0, a.slots, sp, lane);
          uint32_t k = 0;
          if (lane == 0) k = atomicAdd(a.chunk_counter, 1u);
          k = __builtin_amdgcn_readfirstlane(k);
          if (k >= a.n_chunks) {
            exhausted = true;
            if (ns) set_point1 = 0xffffffffu; else set_point0 = 0xffffffffu;
            continue;
          }
          // the chunk's point: the last one whose first chunk is <= k
          // (points without packets share their successor's first chunk
          // and are passed over)
          uint32_t lo = 0, hi = a.n_points;
          while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if (a.ranges[mid].chunk0 <= k) lo = mid; else hi = mid;
          }
          const Range rg = a.ranges[lo];
          const uint64_t off = (uint64_t)(k - rg.chunk0) * a.chunk_packets;
          const uint64_t cnt = rg.n_packets - off < (uint64_t)a.chunk_packets ? rg.n_packets - off
                                                                              : (uint64_t)a.chunk_packets;
          copy_point<R>(&pts[ns], &gpts[lo], lane);
          __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
          __builtin_amdgcn_wave_barrier();
          __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
          if (ns) set_point1 = lo; else set_point0 = lo;
          cur_set = ns;
          win_next = rg.first_pid + off;
          win_end = win_next + cnt;
          win_first = win_next;
          win_outbase = rg.out_base + off;
          win_stream = rg.stream;
        }
        const uint64_t avail = win_end - win_next;
        const uint32_t nneed = (uint32_t)__popcll(need);
        const uint32_t take = avail < nneed ? (uint32_t)avail : nneed;
        const uint32_t rank = lane_rank(need);
        const bool got = phase == PH_NEED && rank < take;
        if (got) {
          pid = win_next + rank;
          set = cur_set;
          stream = win_stream;
          if (kPerPacket) outidx = win_outbase + (pid - win_first);
          phase = PH_ENTER;
        }
        fresh |= __ballot(got);
        win_next += take;
        if (cur_set) outstanding1 += take; else outstanding0 += take;
        need = __ballot(phase == PH_NEED);
      }
      if (__builtin_amdgcn_inverse_ballot_w64(fresh)) {
        pf = philox_fixed((uint32_t)pid, (uint32_t)(pid >> 32), stream, k0, k1);
        w = (R)1;
        muz = point_at<R>(pts, set).mu0_z;
        side = 0;
        A = (R)1;
        L_thr = point_at<R>(pts, set).roul_thr;
        ev = 0;
      }
    }

    // ---------------- flight steps: an interface pass if enough lanes wait
    // at an interface, then one event per lane in flight (and, in the first
    // step after a rare pass, per entering lane); a lane that reaches a
    // sphere or its end stops for this iteration
    // lanes in flight for the step's events: the rare pass leaves the lanes
    // in flight as they were, so the first step starts from this
    // iteration's ballot; a step's ballot is taken again only after an
    // interface pass
    uint64_t fmask = fly;
#pragma unroll
    for (int step = 0; step < FlightSteps<R>::value; ++step) {
    if (step > 0 && near_limit) break;
    const Point<R>& P = point_at<R>(pts, set);
    bool bpass = bpass0;  // decided with this iteration's first ballots
    if (step > 0) {
      const uint64_t f = __ballot(phase == PH_FLIGHT);
      const uint64_t b = __ballot(phase == PH_BOUND);
      // b != 0 && (f == 0 || |b| >= bound_thr), with bound_thr >= 1 (as
      // 32-bit popcounts: scalar instructions only)
      bpass = (uint32_t)(__builtin_popcount((uint32_t)b) + __builtin_popcount((uint32_t)(b >> 32))) >=
              (f == 0 ? 1u : bound_thr);
      fmask = f;
    }
    // ---------------- interface pass: the hits recorded since the last one
    // (CUDAISMLMC.h:180-257), with the draws of the hit's own event
    if (bpass) {
    if (phase == PH_BOUND) {
      const bool up = muz > (R)0;  // the direction the packet meets the interface in
      const R zb = DM<R>::csign(L_half, muz);  // the face met
      const R dz = zb - zr;
      const R d = dz * DM<R>::rcp(muz);
      zr = zb;
      // the step left (in a slide s is never read, and its transmission
      // into the next layer multiplies it by f.mut = 0), at least 0: a hop
      // that ended within rounding of the face can leave s - d (d from an
      // approximate reciprocal) a few ulps below 0, and a reflected packet
      // with a negative step would hit again at once, be taken to the far
      // face and, under total internal reflection at both faces, bounce
      // there until the event limit.  The reference subtracts d only when
      // s >= d (MLMC.py:100-119); a step of 0 is its s == d case, an
      // interaction on the face.
      s = DM<R>::max0(s - d);
      const Iface<R> f = layer_rec<R>(pts, lofs).out[up ? 1 : 0];
      const bool outside = f.mut < (R)0;  // no layer beyond (from the host)
      R ct;
      const R Rf = fresnel<R>(f.eta, f.qq, muz, ct);
      e_bnd += 1;
      phase = PH_FLIGHT;
      // reflect when u <= R: the reference draws u in [0, 1) and reflects
      // when u < R (MLMC.py:122-144); this u is in (0, 1], and with u <= R
      // the reflection probability is R on the draws' grid and total
      // internal reflection (R = 1) reflects always
      constexpr uint32_t kLayerBytes = (uint32_t)sizeof(Layer<R>);
      if (bu0 <= Rf) {
        muz = -muz;
      } else {
        const R muz0 = muz;  // the incident direction
        muz = DM<R>::csign(ct, muz);
        if (outside) {
          side = up ? 1 : 0;
          phase = PH_SPHERE;  // su1, su2 hold the hit's return draw and new |mu_z|
        } else {
          // into the next layer; f.cmin is 0 unless it is a denser glass
          // layer, f.mut is mu_t here / mu_t there where both layers
          // scatter, 0 out of a glass layer (the zero step into a
          // scattering layer, CUDAISMLMC.h:230-235), 1 into one
          uint32_t nofs = up ? lofs + kLayerBytes : lofs - kLayerBytes;
          const R cg = glass_entry_cos<R>(ct, f.cmin);  // the cosine in the next layer
          bool enter = true;
          if (kSlides && f.mut == (R)INFINITY) {  // the next layer is a slide
            // A slide (lossless, plane-parallel) entered through this face
            // (reflectance Rf, from either side by reciprocity) sends the
            // packet on between this face and its far face (r2, at the
            // slide's cosine) until it leaves: back through this face with
            // probability r2 (1 - Rf) / (1 - Rf r2), through the far face
            // otherwise, the geometric series of the bounces MLMC.py:114-163
            // simulates one face at a time.  The hit's draw, rescaled to
            // the event u > Rf, decides.  Either way the packet enters a
            // scattering layer from glass: zero step (CUDAISMLMC.h:230-235).
            const Iface<R> g = layer_rec<R>(pts, nofs).out[up ? 1 : 0];
            R ct3;
            const R r2 = fresnel<R>(g.eta, g.qq, cg, ct3);
            s = (R)0;
            if ((bu0 - Rf) * ((R)1 - Rf * r2) <= r2 * ((R)1 - Rf) * ((R)1 - Rf)) {
              muz = -muz0;  // back into this layer, on the face
              enter = false;
            } else if (g.mut < (R)0) {
              muz = DM<R>::csign(ct3, muz0);  // out of the stack through the slide
              side = up ? 1 : 0;
              phase = PH_SPHERE;
              enter = false;
            } else {
              muz = DM<R>::csign(ct3, muz0);  // into the layer beyond the slide
              nofs = up ? nofs + kLayerBytes : nofs - kLayerBytes;
            }
          } else {
            muz = DM<R>::csign(cg, muz);
            s *= f.mut;
          }
          if (enter) {
            lofs = nofs;
            const Layer<R>& NL = layer_rec<R>(pts, lofs);
            zr = up ? -NL.half : NL.half;  // the near face of the next layer
            L_half = NL.half; L_step = NL.step_k; L_abs = NL.absfrac;
            L_ha = NL.hg_a; L_hb = NL.hg_b; L_hc = NL.hg_c; L_hd = NL.hg_d; L_he = NL.hg_e;
            L_kind = NL.kind;
          }
        }
      }
      // a packet still running after kMaxEventsPerPacket events is dropped
      // (counted, not tallied); only a wave that was near the limit at the
      // top of this iteration can hold such a lane
      if (near_limit && phase == PH_FLIGHT && ev > kMaxEventsPerPacket) {
        phase = PH_END;
        aborted = true;
      }
    }
    fmask = __ballot(phase == PH_FLIGHT);
    }

    const bool run_flight = __builtin_amdgcn_inverse_ballot_w64(fmask);
    const bool run_enter = step == 0 && rare && (phase == PH_ENTER);
    if ((fmask | __ballot(run_enter)) == 0) break;
    uint32_t rw[4];
    philox_block(pf, ev, k0, k1, rw);
    // The event's first three uniforms go straight into the registers an
    // interface hit or a reflected entrance keeps them in (bu0: Fresnel
    // draw; su1, su2: return draw and new |mu_z|).  Every lane that runs
    // an event overwrites them; they are read only after a hit or a
    // reflected entrance, whose own event is the last to write them, and a
    // lane waiting for that read runs no event meanwhile.
    if (run_flight || run_enter) {
      bu0 = u01<R>(rw[0]);
      su1 = Draw<R>::keep(rw[1]);
      su2 = Draw<R>::keep(rw[2]);
    }
    const R u0 = bu0, u1 = su1, u2 = su2, u3 = u01<R>(rw[3]);

    if (run_flight) {
      const bool glass = (L_kind >= KIND_GLASS);
      // The reference hits when s >= d = dz / mu_z, with d infinite for
      // mu_z = 0 (MLMC.py:100-114).  Here the hop's end zr + s mu_z lies
      // beyond the face ahead: s |mu_z| > |dz|, the same test except where
      // the two are exactly equal (an interaction that ends on the
      // interface, whose next event is then the hit), and mu_z = 0 never
      // hits.  Bitwise, not short-circuit: no branches around the
      // comparisons.
      const R zn = s * muz + zr;
      const bool hit = glass | (DM<R>::abs(zn) > L_half);
      // A hit reaches the interface (its draws wait in bu0, su1, su2);
      // otherwise hop, drop, roulette, spin (CUDAISMLMC.h:259-309).  The
      // interaction's new state is computed for every lane in flight (the
      // wave runs this code whenever one of its lanes interacts) and taken
      // where the lane interacted, which saves the copies that merging two
      // branches costs.
      // A packet's weight reaches 0 only where roulette strikes (w - w
      // mu_a/mu_t > 0 in a turbid layer down to the roulette threshold), so
      // the roulette branch, which runs only when a lane's weight falls to
      // the threshold, is where a packet ends; the new direction and step
      // of a packet that ended are never read.
      R w_new = w - w * L_abs;
      phase = hit ? PH_BOUND : PH_FLIGHT;
      if (!hit && w_new <= L_thr) {
        const bool survive = !P.cuda_roulette && (u3 <= P.roul_p);
        const R w_r = survive ? w_new * P.roul_m : (R)0;
        A += w_r - w_new;
        w_new = w_r;
        if (!(w_r > (R)0)) phase = PH_END;
      }
      const R ct = Draw<R>::hg_cos(L_kind, L_ha, L_hb, L_hc, L_hd, L_he, u1);
      const R muz_new = spin_z<R>(muz, ct, u2);
      const R s_new = DM<R>::nlog2(u0) * L_step;
      if (!hit) {
        zr = zn;
        w = w_new;
        muz = muz_new;
        s = s_new;
      }
      ev += 1;
    } else if (run_enter) {
      // (re-)entrance through the outer face (CUDAISMLMC.h:367-369, 502-503)
      e_inc += 1;
      R ct;
      const R Re = entrance_R<R>(P, side, muz, ct);
      // reflected when u <= Re (u in (0, 1]): a totally reflecting entrance
      // (Re = 1) reflects every packet, as at an interface
      if (u0 <= Re) {
        phase = PH_SPHERE;  // su1, su2 hold this event's return draw and new |mu_z|
      } else {
        constexpr uint32_t kLayerBytes = (uint32_t)sizeof(Layer<R>);
        lofs = layer_offset<R>(set, P.first_layer[side]);
        const R cg = glass_entry_cos<R>(ct, P.ent_cmin[side]);  // the cosine in the first layer
        if (kSlides && layer_rec<R>(pts, lofs).kind == KIND_SLIDE) {
          // The first layer is a slide: the packet runs in it between the
          // ambient face (rA) and the face of the layer behind (rB), both at
          // the slide's cosine, until it leaves: back out through the
          // ambient face with probability rB (1 - rA) / (1 - rA rB) (the
          // interface pass's closed form, from inside), the entrance draw
          // rescaled to the event u >= Re deciding; the packet then leaves
          // as a reflected entrance does.  Otherwise it enters the layer
          // behind with a zero step (CUDAISMLMC.h:230-235).
          const Layer<R>& SL = layer_rec<R>(pts, lofs);
          const Iface<R> fa = SL.out[side ? 1 : 0], fb = SL.out[side ? 0 : 1];
          R cta, ctb;
          const R rA = fresnel<R>(fa.eta, fa.qq, cg, cta);
          const R rB = fresnel<R>(fb.eta, fb.qq, cg, ctb);
          if ((u0 - Re) * ((R)1 - rA * rB) < ((R)1 - Re) * rB * ((R)1 - rA)) {
            phase = PH_SPHERE;  // su1, su2 hold this event's return draw and new |mu_z|
          } else {
            muz = DM<R>::csign(ctb, muz);
            lofs = side ? lofs - kLayerBytes : lofs + kLayerBytes;
            const Layer<R>& NL = layer_rec<R>(pts, lofs);
            L_half = NL.half; L_step = NL.step_k; L_abs = NL.absfrac;
            L_ha = NL.hg_a; L_hb = NL.hg_b; L_hc = NL.hg_c; L_hd = NL.hg_d; L_he = NL.hg_e;
            L_kind = NL.kind;
            zr = side ? NL.half : -NL.half;  // the face entered: the layer's bottom / top
            s = (R)0;
            phase = PH_FLIGHT;
          }
        } else {
          const Layer<R>& NL = layer_rec<R>(pts, lofs);
          L_half = NL.half; L_step = NL.step_k; L_abs = NL.absfrac;
          L_ha = NL.hg_a; L_hb = NL.hg_b; L_hc = NL.hg_c; L_hd = NL.hg_d; L_he = NL.hg_e;
          L_kind = NL.kind;
          muz = DM<R>::csign(cg, muz);
          zr = side ? NL.half : -NL.half;  // the stack's face: first layer's bottom / last's top
          s = L_kind >= KIND_GLASS ? (R)0 : DM<R>::nlog2(Draw<R>::u(u1)) * L_step;
          phase = PH_FLIGHT;
        }
      }
      ev += 1;
    }
    // a packet that goes on flying past the event limit is dropped at the
    // top of the next iteration; one waiting at an interface or for its
    // sphere decision is checked when it is known whether it continues
    }
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  if (set_point0 != 0xffffffffu) flush_slot<R>(acc0, a.slots, set_point0, lane);
  if (set_point1 != 0xffffffffu) flush_slot<R>(acc1, a.slots, set_point1, lane);
}

__global__ void zero_u64_kernel(unsigned long long* p, size_t n, uint32_t* counter) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0ull;
  if (i == 0) *counter = 0u;
}

// Device helpers on host inputs (parity aid for ismc_eval_physics).
template <typename R>
__global__ void eval_kernel(int op, const Point<R>* pt, const double* in4, int n, double* out2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double* x = in4 + 4 * i;
  double o0 = 0, o1 = 0;
  if (op == 0) {
    // the interface constants as the host derives them (FP64, then R)
    R ct;
    const double q = (x[1] - x[0]) / (x[1] + x[0]);
    o0 = (double)fresnel<R>((R)(x[0] / x[1]), (R)(q * q), (R)x[2], ct);
    o1 = (double)ct;
  } else if (op == 1) {
    // the scattering cosine as the transport kernel forms it, from the
    // layer constants the host stores for g
    R hc[5];
    Draw<R>::constants(x[0], hc);
    const int kind = x[0] == 0.0 ? KIND_ISO : KIND_HG;
    o0 = (double)Draw<R>::hg_cos(kind, hc[0], hc[1], hc[2], hc[3], hc[4], Draw<R>::from_u(x[1]));
  } else if (op == 2) {
    o0 = (double)DM<R>::cos2pi((R)x[0]);
  } else if (op == 3) {
    o0 = (double)(DM<R>::nlog2((R)x[0]) * (R)0.6931471805599453);
  } else if (op == 4) {
    R ct;
    const R mz = (R)x[0];
    o0 = (double)entrance_R<R>(*pt, mz > (R)0 ? 0 : 1, mz, ct);
    o1 = (double)ct;
  } else if (op == 5) {
    R hc[5];
    Draw<R>::constants(x[1], hc);
    const int kind = x[1] == 0.0 ? KIND_ISO : KIND_HG;
    const R c = Draw<R>::hg_cos(kind, hc[0], hc[1], hc[2], hc[3], hc[4], Draw<R>::from_u(x[2]));
    o0 = (double)spin_z<R>((R)x[0], c, Draw<R>::from_u(x[3]));
    o1 = (double)c;
  }
  out2[2 * i] = o0;
  out2[2 * i + 1] = o1;
}

}  // namespace ismc

// ===========================================================================
// host side
// ===========================================================================
using namespace ismc;

struct DevBuf {
  void* p = nullptr;
  size_t cap = 0;
};

struct ismc_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  hipDeviceProp_t prop{};
  std::string err;
  DevBuf points, ranges, slots, counter, per_packet, eval_in, eval_out;
  void* h_stage = nullptr;  // pinned staging
  size_t h_stage_cap = 0;
  // no-allocation launches: this device's address of the static scratch
  // array, and a fixed pinned staging buffer allocated with the context
  void* scratch = nullptr;
  void* h_fixed = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  float last_ms = 0.f;
  int32_t last_grid = 0;
  int64_t last_chunks = 0;
  int blocks_per_cu = 8;
  int chunk_packets = 4096;
  // measured on MI355X (packets/s): without slides (three layers, one
  // layer with mu_s 10) 10 / 12 beat 8 / 16 and 12 / 16 by 0.5-2 %; with
  // slides (the slab between them, whose slides the kernel takes in closed
  // form: 2.3 interface hits per packet) 12 / 16 beats 10 / 12 by 1.5 %
  int idle_lanes = 10;
  int bound_lanes = 12;
  int idle_lanes_slides = 12;
  int bound_lanes_slides = 16;
  // captured launch sequences, replayed when a call repeats a shape
  struct Graph {
    std::vector<uint64_t> key;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
  };
  std::vector<Graph> graphs;
  std::vector<std::vector<uint64_t>> seen;  // shapes run once, captured when they come back
  int64_t graph_launches = 0;
  int64_t graph_captures = 0;
};

static void drop_graphs(ismc_ctx* c) {
  for (auto& g : c->graphs) {
    if (g.exec) (void)hipGraphExecDestroy(g.exec);
    if (g.graph) (void)hipGraphDestroy(g.graph);
  }
  c->graphs.clear();
  c->seen.clear();
}

static int set_err(ismc_ctx* c, int code, const char* fmt, const char* a = "", long long b = 0) {
  char buf[512];
  snprintf(buf, sizeof(buf), fmt, a, b);
  if (c) c->err = buf;
  return code;
}

#define HIPCHECK(ctx, call)                                                              \
  do {                                                                                   \
    hipError_t e_ = (call);                                                              \
    if (e_ != hipSuccess)                                                                \
      return set_err(ctx, ISMC_EDEVICE, "%s (line %lld)", hipGetErrorString(e_), __LINE__); \
  } while (0)

static void drop_graphs(ismc_ctx* c);

static int ensure_dev(ismc_ctx* c, DevBuf& b, size_t bytes) {
  if (b.cap >= bytes && b.p) return 0;
  drop_graphs(c);  // captured graphs hold the old addresses
  if (b.p) {
    hipError_t e = hipFree(b.p);
    (void)e;
    b.p = nullptr;
    b.cap = 0;
  }
  size_t want = std::max<size_t>(bytes, 256);
  hipError_t e = hipMalloc(&b.p, want);
  if (e != hipSuccess) {
    b.p = nullptr;
    return set_err(c, ISMC_ENOMEM, "hipMalloc(%s) failed for %lld bytes", hipGetErrorString(e),
                   (long long)want);
  }
  b.cap = want;
  return 0;
}

static int ensure_stage(ismc_ctx* c, size_t bytes) {
  if (c->h_stage_cap >= bytes) return 0;
  drop_graphs(c);
  if (c->h_stage) {
    hipError_t e = hipHostFree(c->h_stage);
    (void)e;
    c->h_stage = nullptr;
    c->h_stage_cap = 0;
  }
  size_t want = std::max<size_t>(bytes, 1 << 16);
  hipError_t e = hipHostMalloc(&c->h_stage, want, hipHostMallocDefault);
  if (e != hipSuccess) {
    c->h_stage = nullptr;
    return set_err(c, ISMC_ENOMEM, "hipHostMalloc(%s) failed for %lld bytes", hipGetErrorString(e),
                   (long long)want);
  }
  c->h_stage_cap = want;
  return 0;
}

static bool finite_pos(double v) { return std::isfinite(v) && v > 0; }

static int validate(ismc_ctx* c, const ismc_config* cfg) {
  if (!cfg) return set_err(c, ISMC_EINVAL, "null config%s", "");
  if (cfg->n_layers < 1 || cfg->n_layers > ISMC_MAX_LAYERS)
    return set_err(c, ISMC_EINVAL, "n_layers must be in 1..16 (got %s%lld)", "", cfg->n_layers);
  if (!finite_pos(cfg->n_ambient)) return set_err(c, ISMC_EINVAL, "n_ambient must be > 0%s");
  if (!(cfg->roulette_threshold >= 0) || !(cfg->roulette_m >= 1))
    return set_err(c, ISMC_EINVAL, "roulette_threshold >= 0 and roulette_m >= 1 required%s");
  if (cfg->roulette_threshold * cfg->roulette_m > 1.0)
    return set_err(c, ISMC_ERANGE, "roulette_threshold * roulette_m must be <= 1%s");
  int turbid = 0;
  for (int i = 0; i < cfg->n_layers; ++i) {
    if (!finite_pos(cfg->n[i]) || !finite_pos(cfg->t[i]))
      return set_err(c, ISMC_EINVAL, "layer %s%lld: n and t must be finite and > 0", "", i);
    if (!(cfg->mu_a[i] >= 0) || !(cfg->mu_s[i] >= 0) || !std::isfinite(cfg->mu_a[i]) ||
        !std::isfinite(cfg->mu_s[i]))
      return set_err(c, ISMC_EINVAL, "layer %s%lld: mu_a, mu_s must be finite and >= 0", "", i);
    if (!(cfg->g[i] > -1.0 && cfg->g[i] < 1.0) && cfg->mu_s[i] != 0.0)
      return set_err(c, ISMC_EINVAL, "layer %s%lld: g must be in (-1, 1)", "", i);
    if (cfg->mu_s[i] != 0.0) {
      if (!finite_pos(cfg->mu_t[i]))
        return set_err(c, ISMC_EINVAL, "layer %s%lld: mu_t must be > 0", "", i);
      ++turbid;
    }
  }
  if (!turbid && cfg->n_layers == 1 && cfg->mu_s[0] == 0.0)
    return set_err(c, ISMC_EINVAL, "a sample needs at least one scattering layer%s");
  const ismc_sphere* sp[2] = {&cfg->R, &cfg->T};
  for (int k = 0; k < 2; ++k) {
    if (sp[k]->present) {
      const double den = 1.0 - sp[k]->pw * (1.0 - sp[k]->f);
      if (!(sp[k]->fp >= 0 && sp[k]->fp <= 1) || !(den > 0) || !std::isfinite(sp[k]->pw) ||
          !std::isfinite(sp[k]->fs))
        return set_err(c, ISMC_EINVAL, "sphere %s: need 0 <= fp <= 1 and pw (1-f) < 1",
                       k ? "T" : "R");
    } else if (!std::isfinite(sp[k]->angle)) {
      return set_err(c, ISMC_EINVAL, "sphere %s: angle must be finite", k ? "T" : "R");
    }
  }
  return 0;
}

// Constants of the interface from index na (a layer of extinction ta, or
// the ambient medium) to index nb (a layer of extinction tb, a glass layer,
// or the ambient medium), in FP64 (the same expressions and order as the
// oracle's, so the FP64 kernel draws its decisions from equal values).
// n_guard: for a glass layer b, its exit index (glass_guard_index).
template <typename R>
static Iface<R> make_iface(double na, double nb, bool a_turbid, double ta, bool b_turbid,
                           bool b_glass, double tb, double n_guard) {
  Iface<R> f;
  f.eta = (R)(na / nb);
  const double q = (nb - na) / (nb + na);
  f.qq = (R)(q * q);
  double cmin = 0.0;
  if (b_glass && nb > n_guard) {
    const double r = n_guard / nb;
    cmin = std::sqrt(std::max(0.0, 1.0 - (1.0 - (double)GlassGuard<R>::value) * r * r));
  }
  f.cmin = (R)cmin;
  // rescales the remaining step on transmission: mu_t(a) / mu_t(b) between
  // two scattering layers, 0 out of a slide (its step is never read; the
  // first step in the layer behind is the zero step), 1 into a slide
  f.mut = a_turbid ? (b_turbid ? (R)(ta / tb) : (R)1) : (R)0;
  return f;
}

// The index that bounds the refracted direction in glass layer g
// (GlassGuard): on each side, the least index from g's neighbour out to
// the medium that bounds g's run of adjacent glass layers (a scattering
// layer, or the ambient medium past the stack's face), that medium
// included; the larger of the two.  A packet with n^2 (1 - mu_z^2) below
// its square crosses every face on that side and leaves the run there.
static double glass_guard_index(const ismc_config* cfg, int g) {
  double side[2];
  for (int s = 0; s < 2; ++s) {
    const int step = s ? 1 : -1;
    double m = INFINITY;
    int j = g + step;
    for (; j >= 0 && j < cfg->n_layers; j += step) {
      m = std::min(m, cfg->n[j]);
      if (cfg->mu_s[j] != 0.0) break;  // the scattering layer that bounds the run
    }
    if (j < 0 || j >= cfg->n_layers) m = std::min(m, cfg->n_ambient);  // the run reaches the stack's face
    side[s] = m;
  }
  return std::max(side[0], side[1]);
}

template <typename R>
static void compile_point(const ismc_config* cfg, uint32_t flags, Point<R>* P) {
  memset(P, 0, sizeof(*P));
  const int L = cfg->n_layers;
  P->n_layers = L;
  P->cuda_roulette = (flags & ISMC_FLAG_CUDA_ROULETTE) ? 1 : 0;
  double z = 0;
  for (int i = 0; i < L; ++i) {
    Layer<R>& Ly = P->layer[i];
    const double mu_t = cfg->mu_t[i], g = cfg->g[i];
    Ly.zc = (R)(z + 0.5 * cfg->t[i]);
    Ly.half = (R)(0.5 * cfg->t[i]);
    z += cfg->t[i];
    Ly.n = (R)cfg->n[i];
    if (cfg->mu_s[i] == 0.0) {
      Ly.kind = KIND_GLASS;
    } else {
      Ly.kind = g == 0.0 ? KIND_ISO : KIND_HG;
      Ly.step_k = (R)(std::log(2.0) / mu_t);
      Ly.absfrac = (R)(cfg->mu_a[i] / mu_t);
      R hc[5];
      Draw<R>::constants(g, hc);
      Ly.hg_a = hc[0];
      Ly.hg_b = hc[1];
      Ly.hg_c = hc[2];
      Ly.hg_d = hc[3];
      Ly.hg_e = hc[4];
      Ly.mu_t = (R)mu_t;
    }
  }
  // slides: glass layers between scattering layers or the ambient medium,
  // whose bounces the kernel takes in closed form (unless the caller asks
  // for the reference's face-by-face simulation)
  if (!(flags & ISMC_FLAG_SLIDE_BOUNCES)) {
    for (int i = 0; i < L; ++i) {
      const bool lo_ok = i == 0 || cfg->mu_s[i - 1] != 0.0;
      const bool hi_ok = i == L - 1 || cfg->mu_s[i + 1] != 0.0;
      if (cfg->mu_s[i] == 0.0 && lo_ok && hi_ok) P->layer[i].kind = KIND_SLIDE;
    }
  }
  // (the faces that lead into a slide are marked below, mut = +inf)
  const double namb = cfg->n_ambient;
  for (int i = 0; i < L; ++i) {
    const bool turbid = cfg->mu_s[i] != 0.0;
    for (int dir = 0; dir < 2; ++dir) {
      const int j = dir ? i + 1 : i - 1;
      const bool inside = j >= 0 && j < L;
      const bool j_glass = inside && cfg->mu_s[j] == 0.0;
      P->layer[i].out[dir] = make_iface<R>(cfg->n[i], inside ? cfg->n[j] : namb, turbid, cfg->mu_t[i],
                                           inside && cfg->mu_s[j] != 0.0, j_glass, inside ? cfg->mu_t[j] : 1.0,
                                           j_glass ? glass_guard_index(cfg, j) : 0.0);
      if (!inside) P->layer[i].out[dir].mut = (R)-1;
      else if (turbid && P->layer[j].kind == KIND_SLIDE) P->layer[i].out[dir].mut = (R)INFINITY;
    }
  }
  P->t_tot = (R)z;
  P->n_amb = (R)cfg->n_ambient;
  P->roul_thr = (R)cfg->roulette_threshold;
  P->roul_m = (R)cfg->roulette_m;
  P->roul_p = (R)(1.0 / cfg->roulette_m);
  P->mu0_z = cfg->specular_included ? (R)0.99 : (R)1;
  const int first[2] = {0, L - 1};
  const int behind[2] = {1, L - 2};
  const ismc_sphere* sp[2] = {&cfg->R, &cfg->T};
  for (int k = 0; k < 2; ++k) {
    P->first_layer[k] = first[k];
    P->ent_glass[k] = cfg->mu_s[first[k]] == 0.0;
    {
      const double n2 = cfg->n[first[k]];
      const double n3 = (behind[k] >= 0 && behind[k] < L) ? cfg->n[behind[k]] : namb;
      const bool slide = cfg->mu_s[first[k]] == 0.0;
      const Iface<R> e1 = make_iface<R>(namb, n2, false, 1.0, false, slide, 1.0,
                                        slide ? glass_guard_index(cfg, first[k]) : 0.0);
      const Iface<R> e2 = make_iface<R>(n2, n3, false, 1.0, false, false, 1.0, 0.0);
      P->ent_eta1[k] = e1.eta; P->ent_qq1[k] = e1.qq; P->ent_cmin[k] = e1.cmin;
      P->ent_eta2[k] = e2.eta; P->ent_qq2[k] = e2.qq;
    }
    if (sp[k]->present) {
      P->cos_spec[k] = (R)std::sqrt(1.0 - sp[k]->fp);
      P->p_return[k] = (R)((sp[k]->pw * sp[k]->fs) / (1.0 - sp[k]->pw * (1.0 - sp[k]->f)));
    } else {
      const double fp = 0.5 * (1.0 - std::cos(2.0 * sp[k]->angle));
      P->cos_spec[k] = (R)std::sqrt(1.0 - fp);
      P->p_return[k] = (R)-1;
    }
  }
}

// The order in which a launch's points hand out their chunks: costliest
// packets first.  Waves take chunks in point order, so a chunk of a
// long-history point that starts last holds the whole launch until it
// ends; started first, it overlaps the rest.  Measured on MI355X: 1024
// points log-uniform in mu_a 0.01-10 and mu_s 1-1000 /cm (0.1 cm, g 0.8),
// 2^17 packets each, 13-15 % less kernel time than in input order; a 32 x 32
// lookup grid at 2^16 packets per point, 19-27 % less.  Only the order
// matters (each point keeps its own Philox stream), and the estimate
// (a diffusion-like count of interactions per packet: mu_s t (1 + mu_s' t)
// damped by absorption, summed over the scattering layers) needs only to
// rank points roughly.  Tallies do not depend on the order.
static double cost_estimate(const ismc_config* cfg) {
  double e = 1.0;
  for (int i = 0; i < cfg->n_layers; ++i) {
    if (cfg->mu_s[i] == 0.0) continue;
    const double tau = cfg->mu_s[i] * cfg->t[i];
    const double tau_r = tau * (1.0 - cfg->g[i]);
    e += tau * (1.0 + tau_r) / (1.0 + 10.0 * cfg->mu_a[i] * cfg->t[i]);
  }
  return e;
}

static std::vector<int> launch_order(const ismc_config* cfgs, int32_t n_points, bool keep) {
  std::vector<int> order((size_t)n_points);
  for (int p = 0; p < n_points; ++p) order[(size_t)p] = p;
  if (keep || n_points < 2) return order;
  std::vector<double> est((size_t)n_points);
  for (int p = 0; p < n_points; ++p) est[(size_t)p] = cost_estimate(&cfgs[p]);
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return est[(size_t)a] > est[(size_t)b]; });
  return order;
}

// Whether a configuration has a slide whose bounces the kernel takes in
// closed form (compile_point marks the same layers).
static bool has_slides(const ismc_config* cfg, uint32_t flags) {
  if (flags & ISMC_FLAG_SLIDE_BOUNCES) return false;
  const int L = cfg->n_layers;
  for (int i = 0; i < L; ++i) {
    const bool lo_ok = i == 0 || cfg->mu_s[i - 1] != 0.0;
    const bool hi_ok = i == L - 1 || cfg->mu_s[i + 1] != 0.0;
    if (cfg->mu_s[i] == 0.0 && lo_ok && hi_ok) return true;
  }
  return false;
}

template <typename R>
static const void* transport_fn(bool per_packet, bool slides) {
  if (per_packet)
    return slides ? (const void*)transport_kernel<R, true, true> : (const void*)transport_kernel<R, true, false>;
  return slides ? (const void*)transport_kernel<R, false, true> : (const void*)transport_kernel<R, false, false>;
}

template <typename R>
static void launch_transport(bool per_packet, bool slides, int grid, hipStream_t st, const Args& a) {
  if (per_packet) {
    if (slides) transport_kernel<R, true, true><<<grid, kBlock, 0, st>>>(a);
    else transport_kernel<R, true, false><<<grid, kBlock, 0, st>>>(a);
  } else {
    if (slides) transport_kernel<R, false, true><<<grid, kBlock, 0, st>>>(a);
    else transport_kernel<R, false, false><<<grid, kBlock, 0, st>>>(a);
  }
}

// Launch shape of a transport launch: workgroups per CU (the occupancy
// limit unless the caller asks for a number) and packets per queue chunk.
template <typename R>
static int launch_shape(ismc_ctx* c, const uint64_t* npk, int32_t n_points, const ismc_tune* tune,
                        bool per_packet, bool slides, int* per_cu_out, int* chunk_out) {
  const int bpc = (tune && tune->blocks_per_cu > 0) ? tune->blocks_per_cu : c->blocks_per_cu;
  // persistent grid: as many blocks as stay resident (occupancy), unless the
  // caller asks for a given number per CU
  int resident = 0;
  HIPCHECK(c, hipOccupancyMaxActiveBlocksPerMultiprocessor(&resident, transport_fn<R>(per_packet, slides),
                                                           kBlock, 0));
  const int per_cu = (tune && tune->blocks_per_cu > 0) ? bpc : std::max(1, resident);
  // packets per queue chunk: the caller's, else a size fitted to the
  // launch.  Each chunk costs a queue atomic, a point-table copy and a
  // tally flush (16 atomics on the point's slots, contended when every
  // wave runs the same point), so small chunks are slow; a wave runs a
  // chunk's packets itself, so large chunks leave a tail of waves
  // finishing long histories.  Measured on MI355X (single points of 1e6,
  // 1e7 and 1e8 packets, 2^26 packets of three layers, 2^28 of the slab
  // between slides, 4096 points of 1e5), the best size grows like the
  // square root of the packets per resident wave: 256 at ~140, 512 at
  // ~1,400, 1024 from ~10^4.  So 14 sqrt(packets per wave), kept within
  // 256 .. 1024 (and the context's default as a ceiling).  A launch of
  // many points switches point at almost every chunk, each switch a flush
  // and a table copy, and does best with 2048 (4096 points of 1e5: 100.6 ms
  // against 105.1 at 1024), so from 256 points the ceiling is 2048.
  int chunk = c->chunk_packets;
  if (tune && tune->chunk_packets > 0) {
    chunk = tune->chunk_packets;
  } else {
    uint64_t requested = 0;
    for (int p = 0; p < n_points; ++p) requested += std::min<uint64_t>(npk[p], kMaxPacketsPerPoint);
    const uint64_t waves = (uint64_t)c->prop.multiProcessorCount * per_cu * kWavesPerBlock;
    const uint64_t fit = 14 * (uint64_t)std::sqrt((double)(requested / waves));
    const uint64_t ceiling = n_points >= 256 ? 2048 : 1024;
    chunk = (int)std::min<uint64_t>((uint64_t)c->chunk_packets,
                                    std::min<uint64_t>(ceiling, std::max<uint64_t>(256, fit)));
  }
  *per_cu_out = per_cu;
  *chunk_out = chunk;
  return 0;
}

// The queue record of point p (chunk numbers run point after point) after
// the range checks; *n_chunks and *total advance past it.
static int point_range(ismc_ctx* c, int p, const uint64_t* npk, const uint64_t* first,
                       const uint32_t* streams, int chunk, uint64_t* n_chunks, uint64_t* total,
                       Range* r) {
  if (npk[p] > kMaxPacketsPerPoint)
    return set_err(c, ISMC_ERANGE, "point %s%lld: more than 2^31 packets in one call", "", p);
  const uint64_t f0 = first ? first[p] : 0ull;
  if (f0 + npk[p] < f0) return set_err(c, ISMC_ERANGE, "packet id range overflows%s");
  if (*n_chunks > 0xffffffffull) return set_err(c, ISMC_ERANGE, "too many chunks%s");
  r->first_pid = f0;
  r->out_base = *total;
  r->n_packets = npk[p];
  r->chunk0 = (uint32_t)*n_chunks;
  r->stream = streams ? streams[p] : (uint32_t)p;
  *n_chunks += (npk[p] + (uint64_t)chunk - 1) / (uint64_t)chunk;
  *total += npk[p];
  return 0;
}

// Kernel arguments of a launch whose buffers are given.
static Args make_args(ismc_ctx* c, const ismc_tune* tune, bool slides, void* points, void* ranges, void* slots,
                      void* counter, double* per_packet, int32_t n_points, uint64_t n_chunks,
                      int chunk, uint64_t seed) {
  Args a{};
  a.points = points;
  a.ranges = (const Range*)ranges;
  a.n_points = (uint32_t)n_points;
  a.n_chunks = (uint32_t)n_chunks;
  a.chunk_packets = (uint32_t)chunk;
  a.seed_lo = (uint32_t)seed;
  a.seed_hi = (uint32_t)(seed >> 32);
  a.chunk_counter = (uint32_t*)counter;
  a.slots = (unsigned long long*)slots;
  a.per_packet = per_packet;
  const int idle0 = slides ? c->idle_lanes_slides : c->idle_lanes;
  const int bound0 = slides ? c->bound_lanes_slides : c->bound_lanes;
  a.idle_lanes = (uint32_t)std::min(64, std::max(1, (tune && tune->idle_lanes > 0) ? tune->idle_lanes : idle0));
  a.bound_lanes = (uint32_t)std::min(64, std::max(1, (tune && tune->bound_lanes > 0) ? tune->bound_lanes : bound0));
  return a;
}

// persistent grid: resident blocks, but no more than the packets fill
static int grid_for(ismc_ctx* c, int per_cu, uint64_t total) {
  const uint64_t waves_useful = std::max<uint64_t>(1, (total + 63) / 64);
  const uint64_t blocks_useful = (waves_useful + kWavesPerBlock - 1) / kWavesPerBlock;
  return (int)std::max<uint64_t>(
      1, std::min<uint64_t>((uint64_t)c->prop.multiProcessorCount * per_cu, blocks_useful));
}

// tallies of one point from its 16 fixed-point slot words
static void slot_to_tally(const unsigned long long* v, uint64_t n_packets, ismc_tally* t) {
  const double inv = 1.0 / 4294967296.0;
  for (int k = 0; k < 5; ++k) {
    t->sum[k] = (double)v[k] * inv;
    t->sumsq[k] = (double)v[5 + k] * inv;
    t->events[k] = v[10 + k];
  }
  t->n_aborted = v[15];
  t->n_packets = n_packets;
}

template <typename R>
static int run_points(ismc_ctx* c, const ismc_config* cfgs, const uint64_t* npk,
                      const uint64_t* first, const uint32_t* streams, int32_t n_points,
                      uint64_t seed, uint32_t flags, const ismc_tune* tune, ismc_tally* out,
                      double* per_packet_host) {
  int per_cu = 1, chunk = 1, rc;
  bool slides = false;
  for (int p = 0; p < n_points && !slides; ++p) slides = has_slides(&cfgs[p], flags);
  if ((rc = launch_shape<R>(c, npk, n_points, tune, per_packet_host != nullptr, slides, &per_cu, &chunk)))
    return rc;
  // points in launch order (costliest first; a per-packet launch keeps its
  // order, its output rows follow the points)
  const std::vector<int> order = launch_order(cfgs, n_points, per_packet_host != nullptr);
  std::vector<Range> ranges((size_t)n_points);
  uint64_t total = 0, n_chunks = 0;
  for (int k = 0; k < n_points; ++k)
    if ((rc = point_range(c, order[(size_t)k], npk, first, streams, chunk, &n_chunks, &total, &ranges[(size_t)k])))
      return rc;
  if (n_chunks > 0xffffffffull) return set_err(c, ISMC_ERANGE, "too many chunks%s");
  const size_t pbytes = sizeof(Point<R>) * (size_t)n_points;
  const size_t cbytes = sizeof(Range) * (size_t)n_points;
  const size_t sbytes = sizeof(unsigned long long) * kSlotWords * (size_t)n_points;
  if ((rc = ensure_dev(c, c->points, pbytes))) return rc;
  if ((rc = ensure_dev(c, c->ranges, cbytes))) return rc;
  if ((rc = ensure_dev(c, c->slots, sbytes))) return rc;
  if ((rc = ensure_dev(c, c->counter, 64))) return rc;
  if ((rc = ensure_stage(c, pbytes + cbytes + sbytes))) return rc;
  char* st = (char*)c->h_stage;
  Point<R>* hp = reinterpret_cast<Point<R>*>(st);
  for (int k = 0; k < n_points; ++k) compile_point<R>(&cfgs[order[(size_t)k]], flags, &hp[k]);
  memcpy(st + pbytes, ranges.data(), cbytes);
  const size_t nslot = (size_t)kSlotWords * n_points;
  if (per_packet_host) {
    if ((rc = ensure_dev(c, c->per_packet, sizeof(double) * 6 * std::max<uint64_t>(total, 1))))
      return rc;
  }
  const Args a = make_args(c, tune, slides, c->points.p, c->ranges.p, c->slots.p, c->counter.p,
                           per_packet_host ? (double*)c->per_packet.p : nullptr, n_points, n_chunks,
                           chunk, seed);
  const int grid = grid_for(c, per_cu, total);
  unsigned long long* hs = reinterpret_cast<unsigned long long*>(st + pbytes + cbytes);
  // the launch sequence: upload point tables and the range records from pinned
  // staging, clear the tally slots and queue counter with a kernel (no
  // memset nodes), run the transport, download the slots
  auto enqueue = [&](bool timed) -> int {
    HIPCHECK(c, hipMemcpyAsync(c->points.p, st, pbytes, hipMemcpyHostToDevice, c->stream));
    HIPCHECK(c, hipMemcpyAsync(c->ranges.p, st + pbytes, cbytes, hipMemcpyHostToDevice, c->stream));
    zero_u64_kernel<<<(unsigned)((nslot + 255) / 256), 256, 0, c->stream>>>(
        (unsigned long long*)c->slots.p, nslot, (uint32_t*)c->counter.p);
    HIPCHECK(c, hipGetLastError());
    if (timed) HIPCHECK(c, hipEventRecord(c->e0, c->stream));
    launch_transport<R>(per_packet_host != nullptr, slides, grid, c->stream, a);
    HIPCHECK(c, hipGetLastError());
    if (timed) HIPCHECK(c, hipEventRecord(c->e1, c->stream));
    HIPCHECK(c, hipMemcpyAsync(hs, c->slots.p, sbytes, hipMemcpyDeviceToHost, c->stream));
    return 0;
  };
  const bool use_graph = tune && tune->use_graph && !per_packet_host;
  ismc_ctx::Graph* hit = nullptr;
  if (use_graph) {
    const std::vector<uint64_t> key = {
        sizeof(R), (uint64_t)n_points, n_chunks, (uint64_t)chunk, (uint64_t)grid,
        (uint64_t)a.idle_lanes, (uint64_t)a.bound_lanes,
        (uint64_t)(uintptr_t)c->points.p, (uint64_t)(uintptr_t)c->ranges.p,
        (uint64_t)(uintptr_t)c->slots.p, (uint64_t)(uintptr_t)c->counter.p,
        (uint64_t)(uintptr_t)c->h_stage, pbytes, cbytes, sbytes, seed, (uint64_t)slides};
    for (auto& g : c->graphs)
      if (g.key == key) hit = &g;
    // a shape is captured the second time it comes: one-off shapes (a
    // solver round with a new number of pending points) run eagerly
    bool capture = false;
    if (!hit) {
      auto it = std::find(c->seen.begin(), c->seen.end(), key);
      capture = it != c->seen.end();
      if (capture) {
        c->seen.erase(it);
      } else {
        if (c->seen.size() >= 32) c->seen.erase(c->seen.begin());
        c->seen.push_back(key);
      }
    }
    if (capture) {
      if (c->graphs.size() >= 8) {  // evict the oldest
        if (c->graphs.front().exec) (void)hipGraphExecDestroy(c->graphs.front().exec);
        if (c->graphs.front().graph) (void)hipGraphDestroy(c->graphs.front().graph);
        c->graphs.erase(c->graphs.begin());
      }
      ismc_ctx::Graph g;
      g.key = key;
      HIPCHECK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
      const int erc = enqueue(false);
      hipGraph_t graph = nullptr;
      const hipError_t ee = hipStreamEndCapture(c->stream, &graph);
      if (erc) {
        if (graph) (void)hipGraphDestroy(graph);
        return erc;
      }
      if (ee != hipSuccess)
        return set_err(c, ISMC_EDEVICE, "hipStreamEndCapture: %s", hipGetErrorString(ee));
      g.graph = graph;
      HIPCHECK(c, hipGraphInstantiate(&g.exec, g.graph, nullptr, nullptr, 0));
      c->graphs.push_back(g);
      hit = &c->graphs.back();
      c->graph_captures++;
    }
  }
  if (use_graph && hit) {
    // the graph's own duration (uploads + clear + transport + download)
    HIPCHECK(c, hipEventRecord(c->e0, c->stream));
    HIPCHECK(c, hipGraphLaunch(hit->exec, c->stream));
    HIPCHECK(c, hipEventRecord(c->e1, c->stream));
    c->graph_launches++;
  } else {
    if ((rc = enqueue(true))) return rc;
  }
  if (per_packet_host)
    HIPCHECK(c, hipMemcpyAsync(per_packet_host, c->per_packet.p, sizeof(double) * 6 * total,
                               hipMemcpyDeviceToHost, c->stream));
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  HIPCHECK(c, hipEventElapsedTime(&c->last_ms, c->e0, c->e1));
  c->last_grid = grid;
  c->last_chunks = (int64_t)n_chunks;
  for (int k = 0; k < n_points; ++k) {
    const int p = order[(size_t)k];
    slot_to_tally(hs + (size_t)k * kSlotWords, npk[p], &out[p]);
  }
  return 0;
}

// ---------------------------------------------------------------------------
// No-allocation launches (ismc_simulate_batch_static).  Every device buffer
// of such a launch (point tables, queue range records, tally slots, queue
// counter) lives in one file-scope device array, `scratch`, which each GPU
// the library runs on holds until the process ends.  In host code the
// array's name only stands for the symbol: the host reaches the memory
// through the address hipGetSymbolAddress returns with the device current
// (kept per device in g_scratch_addr, looked up once per device when a
// context is created), and the kernels take that address as an argument.
// All calls on one GPU share its array and the last call's contents stay
// there: the slots and the counter are cleared by a kernel at every launch,
// and calls that overlap on one GPU race.
constexpr int32_t kStaticMaxPoints = ISMC_STATIC_MAX_POINTS;
constexpr size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }
constexpr size_t kStaticPointBytes = align256(sizeof(Point<double>) * kStaticMaxPoints);
constexpr size_t kStaticRangeBytes = align256(sizeof(Range) * kStaticMaxPoints);
constexpr size_t kStaticSlotBytes = align256(sizeof(unsigned long long) * kSlotWords * kStaticMaxPoints);
constexpr size_t kStaticBytes = kStaticPointBytes + kStaticRangeBytes + kStaticSlotBytes + 256;
static_assert(sizeof(Point<float>) <= sizeof(Point<double>), "FP32 tables fit the FP64 section");
static_assert(kStaticBytes == ISMC_STATIC_SCRATCH_BYTES, "ismc.h states the scratch size");

static __device__ __align__(256) unsigned char scratch[kStaticBytes];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void* g_scratch_addr[kMaxDevices];

// this device's address of `scratch`; the device must be current
static hipError_t scratch_address(int device, void** out) {
  *out = nullptr;
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_addr[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch_addr[device] = p;
  }
  *out = g_scratch_addr[device];
  return hipSuccess;
}

constexpr size_t kFixedStageBytes = 256 * 1024;  // pinned, allocated with the context

template <typename R>
static int run_points_static(ismc_ctx* c, const ismc_config* cfgs, const uint64_t* npk,
                             const uint64_t* first, const uint32_t* streams, int32_t n_points,
                             uint64_t seed, uint32_t flags, const ismc_tune* tune, ismc_tally* out) {
  if (!c->scratch || !c->h_fixed)
    return set_err(c, ISMC_EDEVICE, "no static scratch address for this device%s");
  if (n_points > kStaticMaxPoints)
    return set_err(c, ISMC_ERANGE, "at most %s%lld configurations per no-allocation call", "",
                   kStaticMaxPoints);
  int per_cu = 1, chunk = 1, rc;
  bool slides = false;
  for (int p = 0; p < n_points && !slides; ++p) slides = has_slides(&cfgs[p], flags);
  if ((rc = launch_shape<R>(c, npk, n_points, tune, false, slides, &per_cu, &chunk))) return rc;
  // every range is checked before any device work
  uint64_t total = 0, n_chunks = 0;
  for (int p = 0; p < n_points; ++p) {
    Range r;
    if ((rc = point_range(c, p, npk, first, streams, chunk, &n_chunks, &total, &r))) return rc;
  }
  if (n_chunks > 0xffffffffull) return set_err(c, ISMC_ERANGE, "too many chunks%s");
  char* base = (char*)c->scratch;
  Point<R>* d_points = reinterpret_cast<Point<R>*>(base);
  Range* d_ranges = reinterpret_cast<Range*>(base + kStaticPointBytes);
  unsigned long long* d_slots = reinterpret_cast<unsigned long long*>(base + kStaticPointBytes + kStaticRangeBytes);
  uint32_t* d_counter = reinterpret_cast<uint32_t*>(base + kStaticPointBytes + kStaticRangeBytes + kStaticSlotBytes);
  // point tables and range records through the fixed pinned buffer, a
  // batch at a time (each batch's copy finishes before the buffer is reused)
  char* st = (char*)c->h_fixed;
  const int per_batch = (int)(kFixedStageBytes / (sizeof(Point<R>) + sizeof(Range)));
  uint64_t t_run = 0, ch_run = 0;
  for (int p0 = 0; p0 < n_points; p0 += per_batch) {
    const int nb = std::min(per_batch, n_points - p0);
    Point<R>* hp = reinterpret_cast<Point<R>*>(st);
    Range* hr = reinterpret_cast<Range*>(st + sizeof(Point<R>) * (size_t)nb);
    for (int i = 0; i < nb; ++i) {
      compile_point<R>(&cfgs[p0 + i], flags, &hp[i]);
      if ((rc = point_range(c, p0 + i, npk, first, streams, chunk, &ch_run, &t_run, &hr[i]))) return rc;
    }
    HIPCHECK(c, hipMemcpyAsync(d_points + p0, hp, sizeof(Point<R>) * (size_t)nb, hipMemcpyHostToDevice, c->stream));
    HIPCHECK(c, hipMemcpyAsync(d_ranges + p0, hr, sizeof(Range) * (size_t)nb, hipMemcpyHostToDevice, c->stream));
    HIPCHECK(c, hipStreamSynchronize(c->stream));
  }
  const size_t nslot = (size_t)kSlotWords * n_points;
  const Args a = make_args(c, tune, slides, d_points, d_ranges, d_slots, d_counter, nullptr, n_points, n_chunks,
                           chunk, seed);
  const int grid = grid_for(c, per_cu, total);
  zero_u64_kernel<<<(unsigned)((nslot + 255) / 256), 256, 0, c->stream>>>(d_slots, nslot, d_counter);
  HIPCHECK(c, hipGetLastError());
  HIPCHECK(c, hipEventRecord(c->e0, c->stream));
  launch_transport<R>(false, slides, grid, c->stream, a);
  HIPCHECK(c, hipGetLastError());
  HIPCHECK(c, hipEventRecord(c->e1, c->stream));
  // tallies back through the same buffer
  const int slot_batch = (int)(kFixedStageBytes / (sizeof(unsigned long long) * kSlotWords));
  const unsigned long long* hs = reinterpret_cast<const unsigned long long*>(st);
  for (int p0 = 0; p0 < n_points; p0 += slot_batch) {
    const int nb = std::min(slot_batch, n_points - p0);
    HIPCHECK(c, hipMemcpyAsync(st, d_slots + (size_t)p0 * kSlotWords,
                               sizeof(unsigned long long) * kSlotWords * (size_t)nb, hipMemcpyDeviceToHost,
                               c->stream));
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    for (int i = 0; i < nb; ++i) slot_to_tally(hs + (size_t)i * kSlotWords, npk[p0 + i], &out[p0 + i]);
  }
  HIPCHECK(c, hipEventElapsedTime(&c->last_ms, c->e0, c->e1));
  c->last_grid = grid;
  c->last_chunks = (int64_t)n_chunks;
  return 0;
}

static int run_dispatch(ismc_ctx* c, const ismc_config* cfgs, const uint64_t* npk,
                        const uint64_t* first, const uint32_t* streams, int32_t n_points,
                        uint64_t seed, uint32_t flags, const ismc_tune* tune, ismc_tally* out,
                        double* per_packet_host, bool no_alloc = false) {
  if (!c) return ISMC_EINVAL;
  if (!cfgs || !npk || !out || n_points < 1) return set_err(c, ISMC_EINVAL, "null argument%s");
  for (int p = 0; p < n_points; ++p) {
    int rc = validate(c, &cfgs[p]);
    if (rc) return rc;
  }
  HIPCHECK(c, hipSetDevice(c->device));
  if (no_alloc) {
    if (flags & ISMC_FLAG_FP64)
      return run_points_static<double>(c, cfgs, npk, first, streams, n_points, seed, flags, tune, out);
    return run_points_static<float>(c, cfgs, npk, first, streams, n_points, seed, flags, tune, out);
  }
  if (flags & ISMC_FLAG_FP64)
    return run_points<double>(c, cfgs, npk, first, streams, n_points, seed, flags, tune, out,
                              per_packet_host);
  return run_points<float>(c, cfgs, npk, first, streams, n_points, seed, flags, tune, out,
                           per_packet_host);
}

extern "C" {

int ismc_create(int device, ismc_ctx** out) {
  if (!out) return ISMC_EINVAL;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return ISMC_EDEVICE;
  if (device < 0 || device >= n) return ISMC_EINVAL;
  ismc_ctx* c = new ismc_ctx();
  c->device = device;
  c->err.reserve(512);  // set_err's messages then never reallocate
  if (hipSetDevice(device) != hipSuccess || hipGetDeviceProperties(&c->prop, device) != hipSuccess ||
      hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&c->e0) != hipSuccess || hipEventCreate(&c->e1) != hipSuccess) {
    ismc_destroy(c);
    return ISMC_EDEVICE;
  }
  // set-up of the no-allocation entry point: this device's address of the
  // static scratch array (once per device) and the fixed pinned staging
  // buffer; without them that entry point reports ISMC_EDEVICE and every
  // other one works as before
  if (scratch_address(device, &c->scratch) != hipSuccess) c->scratch = nullptr;
  if (hipHostMalloc(&c->h_fixed, kFixedStageBytes, hipHostMallocDefault) != hipSuccess) c->h_fixed = nullptr;
  *out = c;
  return 0;
}

void ismc_destroy(ismc_ctx* c) {
  if (!c) return;
  hipError_t e = hipSetDevice(c->device);
  if (c->stream) e = hipStreamSynchronize(c->stream);
  drop_graphs(c);
  DevBuf* bufs[] = {&c->points, &c->ranges, &c->slots, &c->counter, &c->per_packet, &c->eval_in,
                    &c->eval_out};
  for (DevBuf* b : bufs)
    if (b->p) e = hipFree(b->p);
  if (c->h_stage) e = hipHostFree(c->h_stage);
  if (c->h_fixed) e = hipHostFree(c->h_fixed);
  if (c->e0) e = hipEventDestroy(c->e0);
  if (c->e1) e = hipEventDestroy(c->e1);
  if (c->stream) e = hipStreamDestroy(c->stream);
  (void)e;
  delete c;
}

const char* ismc_last_error(const ismc_ctx* c) { return c ? c->err.c_str() : "null context"; }

int ismc_validate_config(const ismc_config* cfg, char* msg, int32_t msg_len) {
  ismc_ctx tmp;
  const int rc = validate(&tmp, cfg);
  if (msg && msg_len > 0) snprintf(msg, (size_t)msg_len, "%s", tmp.err.c_str());
  return rc;
}

int ismc_device_info_get(ismc_ctx* c, ismc_device_info* out) {
  if (!c || !out) return ISMC_EINVAL;
  memset(out, 0, sizeof(*out));
  snprintf(out->name, sizeof(out->name), "%s", c->prop.name);
  snprintf(out->arch, sizeof(out->arch), "%s", c->prop.gcnArchName);
  out->n_cu = c->prop.multiProcessorCount;
  out->clock_khz = c->prop.clockRate;
  out->total_mem = (int64_t)c->prop.totalGlobalMem;
  out->warp_size = c->prop.warpSize;
  return 0;
}

int ismc_simulate(ismc_ctx* c, const ismc_config* cfg, uint64_t n_packets, uint64_t seed,
                  uint64_t first_packet_id, uint32_t stream, uint32_t flags, ismc_tally* out) {
  return run_dispatch(c, cfg, &n_packets, &first_packet_id, &stream, 1, seed, flags, nullptr, out,
                      nullptr);
}

int ismc_simulate_batch(ismc_ctx* c, const ismc_config* cfgs, const uint64_t* n_packets,
                        const uint64_t* first_packet_id, const uint32_t* streams,
                        int32_t n_points, uint64_t seed, uint32_t flags, const ismc_tune* tune,
                        ismc_tally* out) {
  return run_dispatch(c, cfgs, n_packets, first_packet_id, streams, n_points, seed, flags, tune,
                      out, nullptr);
}

int ismc_simulate_batch_static(ismc_ctx* c, const ismc_config* cfgs, const uint64_t* n_packets,
                               const uint64_t* first_packet_id, const uint32_t* streams,
                               int32_t n_points, uint64_t seed, uint32_t flags, const ismc_tune* tune,
                               ismc_tally* out) {
  return run_dispatch(c, cfgs, n_packets, first_packet_id, streams, n_points, seed, flags, tune, out,
                      nullptr, true);
}

int ismc_simulate_packets(ismc_ctx* c, const ismc_config* cfg, uint64_t n_packets, uint64_t seed,
                          uint64_t first_packet_id, uint32_t stream, uint32_t flags,
                          double* per_packet, ismc_tally* out) {
  if (!per_packet) return set_err(c, ISMC_EINVAL, "null per_packet buffer%s");
  if (n_packets > (1ull << 26)) return set_err(c, ISMC_ERANGE, "per-packet output limited to 2^26 packets%s");
  return run_dispatch(c, cfg, &n_packets, &first_packet_id, &stream, 1, seed, flags, nullptr, out,
                      per_packet);
}

int ismc_graph_stats(ismc_ctx* c, int64_t* captures, int64_t* launches) {
  if (!c) return ISMC_EINVAL;
  if (captures) *captures = c->graph_captures;
  if (launches) *launches = c->graph_launches;
  return 0;
}

int ismc_last_launch(ismc_ctx* c, float* kernel_ms, int32_t* grid_blocks, int32_t* block_threads,
                     int64_t* n_chunks) {
  if (!c) return ISMC_EINVAL;
  if (kernel_ms) *kernel_ms = c->last_ms;
  if (grid_blocks) *grid_blocks = c->last_grid;
  if (block_threads) *block_threads = kBlock;
  if (n_chunks) *n_chunks = c->last_chunks;
  return 0;
}

int ismc_eval_physics(ismc_ctx* c, int32_t op, const ismc_config* cfg, const double* in4,
                      int32_t n, uint32_t flags, double* out2) {
  if (!c || !in4 || !out2 || n < 0 || op < 0 || op > 5) return set_err(c, ISMC_EINVAL, "bad eval_physics arguments%s");
  if (op == 4) {
    int rc = validate(c, cfg);
    if (rc) return rc;
  }
  if (n == 0) return 0;
  HIPCHECK(c, hipSetDevice(c->device));
  int rc;
  const bool f64 = flags & ISMC_FLAG_FP64;
  const size_t pbytes = f64 ? sizeof(Point<double>) : sizeof(Point<float>);
  if ((rc = ensure_dev(c, c->eval_in, sizeof(double) * 4 * (size_t)n))) return rc;
  if ((rc = ensure_dev(c, c->eval_out, sizeof(double) * 2 * (size_t)n))) return rc;
  if ((rc = ensure_dev(c, c->points, pbytes))) return rc;
  if ((rc = ensure_stage(c, pbytes))) return rc;
  if (cfg) {
    if (f64) compile_point<double>(cfg, flags, (Point<double>*)c->h_stage);
    else compile_point<float>(cfg, flags, (Point<float>*)c->h_stage);
  } else {
    memset(c->h_stage, 0, pbytes);
  }
  HIPCHECK(c, hipMemcpyAsync(c->points.p, c->h_stage, pbytes, hipMemcpyHostToDevice, c->stream));
  HIPCHECK(c, hipMemcpyAsync(c->eval_in.p, in4, sizeof(double) * 4 * n, hipMemcpyHostToDevice, c->stream));
  const int grid = (n + 255) / 256;
  if (f64)
    eval_kernel<double><<<grid, 256, 0, c->stream>>>(op, (const Point<double>*)c->points.p,
                                                     (const double*)c->eval_in.p, n, (double*)c->eval_out.p);
  else
    eval_kernel<float><<<grid, 256, 0, c->stream>>>(op, (const Point<float>*)c->points.p,
                                                    (const double*)c->eval_in.p, n, (double*)c->eval_out.p);
  HIPCHECK(c, hipGetLastError());
  HIPCHECK(c, hipMemcpyAsync(out2, c->eval_out.p, sizeof(double) * 2 * n, hipMemcpyDeviceToHost, c->stream));
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  return 0;
}

}  // extern "C"
