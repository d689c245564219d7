// psr_internal.h -- device-side data layout shared by the gfx950 kernels.
//
// HBM layout (all per context, particles in octree-key order):
//   ps      float4[npar]   x, y, z, sml            (traversal: 16 B/particle)
//   orig    uint32[npar]   caller index | PSR_INVALID_BIT (PF_INVALID)
//   mass, rho, ie, yeMET          float[npar]
//   lambdaH, lambdaHeI, lambdaHeII double[npar]  (x_HI, x_HeI, x_HeII)
//   yGdepHI/HeI/HeII, yGrec*, heat float[npar]   (f32 atomics)
//   lasthit int64[npar], hits int32[npar], active uint8[npar]
//   nodes   Node[nnode]    32 B: float bot[3], top[3]; int a, b
//           internal: a = first child (8 contiguous), b = -1
//           leaf:     a = first particle, b = particle count (0..8)
//   hit slots (per ray, capacity hit_cap): d f64, b f64, slot u32
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <deque>
#include <vector>
#include <string>

#include "psphray_amd.h"

#define PSR_INVALID_BIT 0x80000000u
#define PSR_MAXDEPTH 21  // octree levels held by a 63-bit key

struct Node {
    float bot[3];
    float top[3];
    int32_t a;
    int32_t b;
};
static_assert(sizeof(Node) == 32, "Node must be 32 bytes");
static_assert(sizeof(psr_ray) == 64, "psr_ray must be 64 bytes");

// tables in device memory (layouts as psr_tables)
struct DevTables {
    const double *ar;  // [ar_nrows][16]: log10 T, then the 15 rate columns
    int32_t ar_nrows;
    double ar_min, ar_step_inv;
    const double *xs;
    int32_t xs_nrows;
    double xs_min, xs_step_inv;
    double xs_cut[3];
    const double *kernel;
    int32_t kernel_n;
    int32_t si_n1, si_n2;
    double si_min1, si_min2, si_step1_inv, si_step2_inv;
    const double *si_x, *si_y;
    const double *si_tab3;  // PhiHI, PhiHeI, Eh interleaved: [si_n1][si_n2][3]
};

// constants the kernels need, passed by value
struct DevPhys {
    double C_H_PER_MASS, C_HE_PER_MASS, C_HI_ENERGY, C_HEI_ENERGY, C_HEII_ENERGY, C_HMF,
        C_HEMF, U_MPROTON_OVER_C_BOLTZMANN, C_BOLTZMANN_OVER_U_MPROTON;
    int32_t isothermal, adiabatic, on_the_spot, h_only, secondary_ionization, no_photon;
    double scaling_fac2_inv, scaling_fac3_inv;
    double fake_temperature;
};

// temperature from internal energy and electrons per hydrogen, and back
// (ieye2T / Tye2ie, src/psystem.h); electrons per hydrogen from the
// fractions (lambda_to_ye, src/psystem.h:134-138)
__device__ __forceinline__ double ieye2T(const DevPhys &P, double ie, double ye) {
    return P.U_MPROTON_OVER_C_BOLTZMANN * ie / (P.C_HEMF * 0.25 + P.C_HMF + P.C_HMF * ye) * 2.0 / 3.0;
}
__device__ __forceinline__ double Tye2ie(const DevPhys &P, double T, double ye) {
    return T * P.C_BOLTZMANN_OVER_U_MPROTON * (P.C_HEMF * 0.25 + P.C_HMF + P.C_HMF * ye) / 2.0 * 3.0;
}
__device__ __forceinline__ double lambda_ye(const DevPhys &P, double lH, double lHeI, double lHeII) {
    return (1.0 - lH) + P.C_HEMF / P.C_HMF * 0.25 * (lHeII + 2 * (1.0 - lHeII - lHeI));
}

// particle state pointers (tree order)
// packed per-particle inputs of the deposit in tree order (sml, mass and the
// fractions the optical depths need), so that a hit gathers one line instead
// of one per array; the chemistry's write-back keeps the fractions current
struct __attribute__((aligned(16))) DepViewH {
    float sml, mass;
    double xHI;
};
struct __attribute__((aligned(32))) DepViewHe {
    float sml, mass;
    double xHI, xHeI, xHeII;
};
static_assert(sizeof(DepViewH) == 16 && sizeof(DepViewHe) == 32, "the deposit reads the view as 16-byte words");

struct DevState {
    const float4 *ps;
    const uint32_t *orig;
    float *mass, *rho, *ie, *yeMET;
    double *lH, *lHeI, *lHeII;
    float *yGdepHI, *yGdepHeI, *yGdepHeII;
    float *yGrecHII, *yGrecHeII, *yGrecHeIII;
    float *heat;
    int64_t *lasthit;
    int32_t *hits;
    uint8_t *active;
    void *dv;        // DepViewHe (dv_he = 1) or DepViewH (0) per particle; -1: not built
    int32_t dv_he;
};

// state, tables and constants as one device-resident block, read by the
// kernels through a pointer (scalar loads on demand instead of a large
// by-value argument list)
struct DevCtx {
    DevState S;
    DevTables T;
    DevPhys P;
};

// hit slots: trace order (d, b, slot), then the per-ray sort writes the
// permutation and the slot / b columns in (d, ipar) order
struct DevHits {
    double *d;
    double *b;
    uint32_t *slot;
    int32_t *count;
    uint32_t *perm;
    uint32_t *sslot;
    double *sb;
    uint32_t *order;  // several rays per workgroup: rays by descending hit count
    int32_t cap;
};

// how a tick's deposit runs: one ray per 3-wavefront workgroup with f32
// atomics or binned records, or two, four or eight rays per 4-wavefront
// workgroup (atomics)
enum { PSR_DEP_WG_ATOMIC = 0, PSR_DEP_WG_BINNED = 1, PSR_DEP_MULTI = 2, PSR_DEP_MULTI2 = 3, PSR_DEP_MULTI8 = 4 };
// rays per workgroup of a deposit mode
inline int psr_dep_rays(int mode) {
    return mode == PSR_DEP_MULTI ? 4 : mode == PSR_DEP_MULTI2 ? 2 : mode == PSR_DEP_MULTI8 ? 8 : 1;
}

// device counters: PSR_NSTAT copies, one 128-byte line each; a block adds
// into copy blockIdx.x % PSR_NSTAT and the host folds the copies.  Device-
// scope atomics on one address serialize beyond the XCD's L2, so a single
// copy costs several ns per wavefront atomic once there are 1e4-1e5 of them.
#define PSR_NSTAT 256
struct __attribute__((aligned(128))) DevStats {
    unsigned long long total_deposit_count;
    unsigned long long saturated_deposit_count;
    unsigned long long active_count;
    unsigned long long solver_steps;
    unsigned long long solver_failures;
    unsigned long long surrogate_count;
    unsigned long long surrogate_fallbacks;
    unsigned long long deposit_processed;
    double lost_photon_count_sum;
    double first_ionization_HI, first_ionization_HeI, first_ionization_HeII;
    double secondary_ionization_HI, secondary_ionization_HeI;
    int32_t flags;  // PSR_DF_* below (copy 0 only)
};
static_assert(sizeof(DevStats) == 128, "DevStats must fill one 128-byte line");

// the counter copy of the calling block
__device__ __forceinline__ DevStats *stat_slot(DevStats *st) { return st + (blockIdx.x & (PSR_NSTAT - 1)); }
#define PSR_DF_HITCAP 1
#define PSR_DF_STACK 2
#define PSR_DF_SORTCAP 4
// a tick whose trace overflowed stops every later kernel until the host has
// replayed it (the flag stays set until then)
#define PSR_DF_HALT (PSR_DF_HITCAP | PSR_DF_STACK)

__device__ __forceinline__ bool halted(const int32_t *halt) { return halt && (*(volatile const int32_t *)halt & PSR_DF_HALT); }

// per-tick values the tick's kernels read from device memory (written by
// k_set_tick at the start of every tick), so that one recorded hipGraph
// serves every tick of the same shape.  Two records alternate between
// consecutive ticks, so that a tick's update can still read its own while
// the next tick's trace runs.
struct DevTick {
    const psr_ray *rays;  // the tick's rays: pool + ray_first
    int64_t tick;
    double tick_time;
    int32_t *slot;  // pinned ring word k_publish writes
    // the flag word as the tick's deposit left it (k_snap_flags): the halt
    // test of its update and the word it publishes, untouched by an
    // overflow in the next tick's trace
    int32_t flags;
};

// a tick launched without waiting for its trace's capacity check; validated
// later from its published flag word (host ring slot)
struct PendingTick {
    int64_t ray_first, nray, tick;
    double tick_time;
    int32_t phases, chem_mode;
    uint32_t slot;
};
#define PSR_RING 256

struct MlpState;
struct EmitState;

struct psr_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    int cu_count = 256;
    std::string arch;  // the device's gcnArchName (tuning signatures)
    // physics
    bool have_physics = false;
    psr_units U{};
    psr_flags F{};
    DevTables T{};
    double *d_tables = nullptr;  // one allocation holding every table
    // tree
    bool have_tree = false;
    int64_t npar = 0;
    int32_t nfield = 0;
    std::vector<double> boxsizes;  // per field
    int64_t nnode = 0;
    int32_t depth = 0;
    int64_t skipped = 0;
    Node *d_nodes = nullptr;
    int32_t *d_node_level = nullptr;
    int32_t *d_field_root = nullptr;
    uint32_t *d_perm = nullptr;  // tree slot -> caller index
    // particles
    bool have_state = false;
    float4 *d_ps = nullptr;
    // the trace's copy of d_ps by leaf: leaf node j's particles at
    // [8 j, 8 j + count), its other slots invalid, so a leaf is one aligned
    // 128-byte line (slots of internal and empty nodes are never read)
    float4 *d_psl = nullptr;
    uint32_t *d_orig = nullptr;
    float *d_f32 = nullptr;   // mass rho ie yeMET yGdep*3 yGrec*3 heat : 11 arrays
    double *d_f64 = nullptr;  // lambdaH lambdaHeI lambdaHeII
    void *d_dview = nullptr;  // the deposit's packed view (32 B per particle)
    bool dview_valid = false;
    int32_t dview_he = -1;
    int64_t *d_lasthit = nullptr;
    int32_t *d_hits = nullptr;
    // particles marked by a tick's trace: two arrays of npar bytes, the
    // tick's parity (par) picks one; the update clears what it used
    uint8_t *d_active = nullptr;
    int par = 0;
    uint8_t *active_of(int p) const { return d_active ? d_active + (size_t)p * (size_t)(npar > 0 ? npar : 1) : nullptr; }
    // rays
    psr_ray *d_rays = nullptr;
    int64_t nray_pool = 0;
    int32_t ray_field_max = -1;   // largest field index in the pool
    EmitState *emit = nullptr;    // device ray emission: sources, spectra
    int64_t ray_cap = 0;
    // hit slots
    DevHits H{};
    int64_t hit_slots = 0;  // number of rays the slot arrays hold
    double *d_ray_len = nullptr, *d_ray_tm = nullptr;
    int64_t last_nray = 0;
    // recombination rays: LTE spectra on the device (energies [3][1024],
    // cdf [3][128][1024], row bounds in T [127]), the rays per species
    // (0: off), threshold and disableLTE, and how many leading rays of the
    // last emitted list make candidates with their hits (all of an uploaded
    // list)
    double *d_lte = nullptr;
    int64_t rec_nrec = 0;
    double rec_threshold = 0.1;
    int32_t rec_disable_lte = 0;
    int64_t rec_cut = INT64_MAX;
    int64_t hits_nray = 0;  // rays whose hit lists hold slots of the current tree (0 after a rebuild)
    // deposit records binned by particle range; d_bin_off holds off[nbins + 1],
    // fill[nbins] and claimed[nbins]
    void *d_bin_rec = nullptr;
    uint32_t *d_bin_off = nullptr;
    int64_t bin_n = 0;
    size_t bin_rec_bytes = 0;
    uint64_t bin_rec_cap = 0;
    int bin_he = 0;
    // stats and phase timers
    DevStats *d_stats = nullptr;
    bool timing = false;
    // event pairs around the trace, sort, deposit and update kernels and the
    // whole tick; ev_mask marks the pairs recorded in the current tick
    hipEvent_t ev[10] = {};
    int ev_mask = 0;
    double t_ms[5] = {0, 0, 0, 0, 0};
    int64_t ticks = 0;
    // chemistry phase 2: compact list of unfinished particles + its length,
    // and their saved integrator state (per particle index)
    uint32_t *d_qlist = nullptr;
    uint32_t *d_q = nullptr;
    double *d_saved = nullptr;
    int64_t qlist_cap = 0;
    // scratch
    void *d_scratch = nullptr;
    size_t scratch_bytes = 0;
    // device copy of DevCtx, refreshed by psr_sync_dev_ctx when dirty
    DevCtx *d_ctx = nullptr;
    bool ctx_dirty = true;
    // surrogate
    MlpState *mlp = nullptr;
    double mlp_residual = 1e-4;
    bool mlp_balance = true;  // also check the ionisation-balance residual
    // ticks launched but not yet validated, and the pinned ring of flag words
    // the device publishes at the end of each tick (0 = not yet written)
    std::deque<PendingTick> pending;
    int32_t *h_ring = nullptr;
    uint32_t seq = 0;
    // deposit accumulation: requested (PSR_ACCUM_*), chosen (ATOMIC or BINNED
    // once decided, AUTO while the trial ticks run), and the trials: ticks
    // alternating the two choices, each with an event pair around its deposit
    static constexpr int TUNE_N = 8;  // tick-graph trial ticks (two pairs per way)
    DevTick *d_tick = nullptr;      // the current tick's record: d_tick_buf + par
    DevTick *d_tick_buf = nullptr;  // two records
    // overlap of a tick's update (on stream2) with the next tick's trace
    // and sort: requested (PSR_TUNE_TICK_OVERLAP), the events between the
    // streams, and whether stream2 holds an update the main stream has not
    // been ordered after yet (psr_join)
    int32_t overlap_req = 1;
    hipStream_t stream2 = nullptr;
    hipEvent_t ev_dep = nullptr, ev_upd = nullptr;
    bool upd_pending = false;
    hipStream_t tick_end = nullptr;  // the stream the last tick ended on
    // an overlapped tick sorts its 8192-hit tier on stream3, beside the
    // 4096-hit tier (their rays are disjoint); the main stream waits for it
    bool sort_split = false;
    hipStream_t stream3 = nullptr;
    hipEvent_t ev_tr = nullptr, ev_s8 = nullptr;
    // the tick recorded as a hipGraph: its executable and the shape it was
    // recorded for (sizes and device pointers its kernels were given)
    int32_t graph_req = -1;  // -1: chosen by timed trials, 0 / 1: off / on
    int32_t graph_use = -1;  // the choice once made
    hipEvent_t gtune_ev[TUNE_N] = {};  // the end of each trial tick
    int32_t gtune_slot[TUNE_N] = {};
    uint32_t gtune_seq[TUNE_N] = {};
    int gtune_n = 0;
    hipGraphExec_t graph_exec = nullptr;
    std::vector<int64_t> graph_key;
    int64_t graph_launches = 0;
    int32_t accum_req = PSR_ACCUM_AUTO;
    int32_t walk_req = PSR_WALK_AUTO;
    int32_t dep_use = -1;            // the decided PSR_DEP_* mode; -1 while timing
    static constexpr int DTUNE_N = 10;  // deposit trials: two per candidate mode
    hipEvent_t tune_ev[2 * DTUNE_N] = {};
    int32_t tune_slot[DTUNE_N] = {};  // ring slot of the trial's tick; -1: a synchronous replay
    int32_t tune_mode[DTUNE_N] = {};
    int tune_n = 0;
    // traversal threads per ray: requested (-1: timed trials, 128 or 256),
    // chosen, and the trials (ticks that trace their rays with both shapes,
    // an event pair around each launch)
    int32_t trace_req = -1;
    int32_t trace_use = -1;
    static constexpr int TTUNE_N = 2;
    hipEvent_t ttune_ev[4 * TTUNE_N] = {};
    int32_t ttune_slot[TTUNE_N] = {};
    int ttune_n = 0;
    bool tune_file_tried = false;  // PSR_TUNING_FILE looked up for this workload

    DevState state() const;
};

// order the main stream after the update of the last overlapped tick (every
// call except psr_tick starts with it; a no-op when nothing is pending)
int psr_join(psr_ctx *c);

// ---- launchers implemented in the .hip files ----
// sync: the trace waits for its capacity check (and retries); otherwise the
// kernels read the halt flag and the tick is validated later
int psr_launch_trace(psr_ctx *c, int64_t ray_first, int64_t nray, bool sync, const DevTick *tk, int nt);
int psr_launch_sort(psr_ctx *c, int64_t nray);
// hit slots for nray rays of cap hits each (reallocates after a sync)
int psr_ensure_hits(psr_ctx *c, int64_t nray, int32_t cap);
int psr_launch_deposit(psr_ctx *c, int64_t ray_first, int64_t nray, int mode);
int psr_launch_ray_order(psr_ctx *c, int64_t nray);
// restart the accumulation trials (tree or physics changed)
void psr_tune_reset(psr_ctx *c);
// timed: record the update phase's event pair (the surrogate records its own);
// the tick and its length come from c->d_tick
int psr_launch_update(psr_ctx *c, bool timed = true);
int psr_launch_update_phase(psr_ctx *c, int phase);
// allocations of the deposit / update phases ahead of a graph recording
int psr_prepare_deposit(psr_ctx *c, int mode);
int psr_prepare_update(psr_ctx *c);
// drop the recorded tick graph (physics, tree, state or options changed)
void psr_graph_reset(psr_ctx *c);
// validate (and if needed replay) every pending tick; blocks
int psr_resolve(psr_ctx *c);
int psr_solve_steps_dev(psr_ctx *c, psr_step *d_steps, int64_t n, double rtol);
int psr_noalloc_init(int device);
void psr_emit_free(psr_ctx *c);
// stable LSD radix sort of (key, value) pairs by key bits [begin, end) and
// an exclusive scan (psr_sort.hip); scratch of the size the *_scratch
// functions give, inputs left unchanged
size_t psr_sort_scratch_u64(int64_t n);
size_t psr_sort_scratch_u16(int64_t n);
size_t psr_sort_scratch_u32(int64_t n);
size_t psr_scan_scratch_i32(int64_t n);
hipError_t psr_sort_pairs_u64(const unsigned long long *kin, const uint32_t *vin, unsigned long long *kout,
                              uint32_t *vout, int64_t n, int begin, int end, void *scratch, hipStream_t st);
hipError_t psr_sort_pairs_u16(const uint16_t *kin, const uint32_t *vin, uint16_t *kout, uint32_t *vout, int64_t n,
                              int begin, int end, void *scratch, hipStream_t st);
hipError_t psr_sort_pairs_u32(const uint32_t *kin, const uint32_t *vin, uint32_t *kout, uint32_t *vout, int64_t n,
                              int begin, int end, void *scratch, hipStream_t st);
hipError_t psr_exclusive_sum_i32(const int32_t *in, int32_t *out, int64_t n, void *scratch, hipStream_t st);
DevPhys psr_dev_phys(const psr_ctx *c);
int psr_sync_dev_ctx(psr_ctx *c);
int psr_ensure_scratch(psr_ctx *c, size_t bytes);
int psr_mlp_update_dev(psr_ctx *c, bool stiff);
// the surrogate's device pointers: parameters, normalisation, and the
// capture buffer of exact steps (cap null when capture is off)
struct MlpDev {
    const float *params;
    const float *norm;
    float *cap;
    unsigned long long *cap_n;
    unsigned long long cap_max;
};
MlpDev psr_mlp_dev(const psr_ctx *c);
// per-particle capture row of the current update (n entries); null on error
uint32_t *psr_mlp_rowof(psr_ctx *c, int64_t n);
bool psr_mlp_ready(const psr_ctx *c);
void psr_mlp_free(psr_ctx *c);

#define PSR_HIP(c, call)                                                           \
    do {                                                                           \
        hipError_t e_ = (call);                                                    \
        if (e_ != hipSuccess) {                                                    \
            (c)->err = std::string(#call) + ": " + hipGetErrorString(e_);          \
            return PSR_ERR_HIP;                                                    \
        }                                                                          \
    } while (0)

// record event pair member (phase k, 0 = start / 1 = end) when timing
#define PSR_EV(c, k, e)                                                          \
    do {                                                                         \
        if ((c)->timing) {                                                       \
            PSR_HIP(c, hipEventRecord((c)->ev[2 * (k) + (e)], (c)->stream));     \
            (c)->ev_mask |= 1 << (k);                                            \
        }                                                                        \
    } while (0)

#define PSR_FAIL(c, code, msg)   \
    do {                         \
        (c)->err = (msg);        \
        return (code);           \
    } while (0)
