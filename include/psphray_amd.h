/*
 * psphray_amd.h -- C ABI of the MI355X (gfx950) photon-packet tracer.
 *
 * Drop-in boundary for the per-tick hot path of P-SPHRAY: the octree build
 * (rt_switch_epoch), ray traversal + per-ray distance sort (rt_trace + trace),
 * photon deposition (deposit), active-set merge (merge_pars) and the
 * per-particle chemistry update (update_pars / step_evolve).  The reference
 * has no FFI layer; its internal C linkage is listed per entry point.
 *
 * Conventions
 *   - every entry point returns PSR_OK (0) or a PSR_ERR_* code and never
 *     signals (the reference raises SIGTRAP via ERROR, support/messages.h:17);
 *     psr_last_error() gives the message;
 *   - all pointers are host pointers borrowed for the duration of the call,
 *     except where a name ends in _dev;
 *   - a context owns its device memory and one HIP stream; one host thread
 *     per context; several contexts may share a device.
 *   - particle arrays are in the caller's order (index ipar, as PSYS(field,
 *     ipar), src/psystem.h:30); the device keeps them in tree order and maps
 *     back on download.
 */
#ifndef PSPHRAY_AMD_H
#define PSPHRAY_AMD_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PSR_OK 0
#define PSR_ERR_ARG 1        /* bad argument / call order */
#define PSR_ERR_HIP 2        /* HIP runtime error */
#define PSR_ERR_CAPACITY 3   /* caller buffer too small; sizes returned */
#define PSR_ERR_TREE 4       /* octree deeper than the key resolution */
#define PSR_ERR_SOLVER 5     /* chemistry did not reach t = 1 */
#define PSR_ERR_STATE 6      /* missing tables / tree / particles */
#define PSR_ERR_NOTSUP 7     /* option outside this implementation */

typedef struct psr_ctx psr_ctx;

/* internal units and constants (src/units.c:66-104; src/config.c:43-51) */
typedef struct psr_units {
    double C_H, C_HMF, C_HEMF, C_1_CMH, C_1_GRAMH, C_1_SECH;
    double U_CM, U_GRAM, U_SEC, U_M, U_KM, U_KPC, U_YR, U_MYR, U_KG, U_MSUN,
           U_MPROTON, U_J, U_EV, U_ERG, U_RY_ENG;
    double C_BOLTZMANN, C_PLANCK, C_SPEED_LIGHT, C_H_PER_MASS, C_HE_PER_MASS,
           C_HI_ENERGY, C_HEI_ENERGY, C_HEII_ENERGY,
           U_MPROTON_OVER_C_BOLTZMANN, C_BOLTZMANN_OVER_U_MPROTON;
} psr_units;

/* run flags: the CFG_* switches of src/config.c:22-42 that the hot path
 * reads.  EOS (psphray.eos) is not supported: PSR_ERR_NOTSUP. */
typedef struct psr_flags {
    int32_t isothermal, adiabatic, on_the_spot, h_only, secondary_ionization,
            trace_only, no_photon, comoving;
    double redshift;
    double fake_temperature;
} psr_flags;

/* host-built lookup tables, all in internal units */
typedef struct psr_tables {
    /* atomic rates (src/atomic-rate.c:253-278): row-major [16][ar_nrows]:
     * logT, then HIci HIIrcA HIIrccA HIIrcB HIIrccB HeIci HeIIci HeIIrcA
     * HeIIrcB HeIIIrcA HeIIIrcB HIcic HIcec Brem Compton */
    int32_t ar_nrows;
    double ar_min, ar_step_inv;
    const double *ar;
    /* cross sections (src/atomic-rate.c:280-318): [4][xs_nrows]: energy,
     * sigma_HI, sigma_HeI, sigma_HeII; xs_cut = 13.6, 24.5, 54.4 eV */
    int32_t xs_nrows;
    double xs_min, xs_step_inv;
    double xs_cut[3];
    const double *xs;
    /* SPH column kernel (src/sph-depth.c:3-33) */
    int32_t kernel_n;
    const double *kernel;
    /* secondary ionization TabFun2 (src/secion.c:75-104): x[si_n1], y[si_n2],
     * then PhiHI, PhiHeI, Eh each [si_n1][si_n2]; si_n1 == 0 disables */
    int32_t si_n1, si_n2;
    double si_min1, si_min2, si_step1_inv, si_step2_inv;
    const double *si_x, *si_y, *si_phi_hi, *si_phi_hei, *si_eh;
} psr_tables;

/* particle system view (PSystem, src/psystem.h:32-78).  For uploads every
 * pointer must be valid; for downloads a NULL pointer skips that field. */
typedef struct psr_state {
    int64_t npar;
    float *pos; /* [npar][3] */
    float *sml, *mass, *rho, *ie, *yeMET;
    int8_t *flag;
    double *lambdaH, *lambdaHeI, *lambdaHeII;
    float *yGdepHI, *yGdepHeI, *yGdepHeII;
    float *yGrecHII, *yGrecHeII, *yGrecHeIII;
    float *heat;
    int64_t *lasthit;
    int32_t *hits;
} psr_state;

/* one ray packet (struct r_t, src/loop.c:21-35), 64 bytes.  field selects
 * the particle field of a batched context (0 for a single field). */
typedef struct psr_ray {
    float s[3];
    float dir[3];
    double Nph;
    double freq;
    double length;
    int32_t type;
    int32_t field;
    int64_t src_or_par;
} psr_ray;

/* chemistry step (Step, src/psystem.h:84-105) */
typedef struct psr_step {
    double lambdaH, lambdaHeI, lambdaHeII, ie, yeMET, yGdepHI, yGdepHeI,
           yGdepHeII, heat, time, nH, T, yGrecHII, yGrecHeII, yGrecHeIII,
           step_remain;
    int32_t refined;
    int32_t nsteps;
    int64_t ipar;
} psr_step;

/* per-tick counters (stat_t subset, src/stat.h:1-56) */
typedef struct psr_stats {
    int64_t ticks;
    int64_t total_deposit_count;      /* sum of full hit-list lengths */
    int64_t saturated_deposit_count;
    int64_t active_count;             /* particles updated */
    int64_t solver_steps;
    int64_t solver_failures;
    int64_t surrogate_count;          /* updates taken from the surrogate */
    int64_t surrogate_fallbacks;      /* surrogate rejected by the residual check */
    int64_t deposit_processed;        /* hits walked by deposit before the cut-off */
    double lost_photon_count_sum;
    double first_ionization_HI, first_ionization_HeI, first_ionization_HeII;
    double secondary_ionization_HI, secondary_ionization_HeI;
} psr_stats;

/* tick phases (src/loop.c:131-148) */
#define PSR_PHASE_TRACE 1
#define PSR_PHASE_DEPOSIT 2
#define PSR_PHASE_UPDATE 4
#define PSR_PHASE_ALL 7
/* chemistry modes: exact; the network for every active particle (the
 * exact solver for the predictions it rejects); or the exact solver's first
 * attempt for every active particle and the network only for the particles
 * that attempt leaves unfinished (the stiff ones), the exact solver
 * continuing for the predictions it rejects */
#define PSR_CHEM_EXACT 0
#define PSR_CHEM_SURROGATE 1
#define PSR_CHEM_SURROGATE_STIFF 2

/* ---- context ---- */
int psr_create(psr_ctx **out, int device);
void psr_destroy(psr_ctx *ctx);
const char *psr_last_error(const psr_ctx *ctx);
int psr_sync(psr_ctx *ctx);

/* ---- host helpers (restating the reference's table set-up) ---- */
/* units_init, src/units.c:66-104 */
int psr_units_init(psr_units *u, double h, double hmf, double hemf, double cmh,
                   double gramh, double sech);
/* builds all tables: ar_init (src/atomic-rate.c:253-274) from an
 * atomic_rates_*.txt file, xs_init (:280-311), the SPH column kernel
 * (src/sph-depth.c) by quadrature, secion_init (src/secion.c:75-99) when
 * with_secion.  The returned struct owns its arrays: psr_tables_free. */
int psr_tables_build(psr_tables *t, const psr_units *u, const char *rates_path,
                     int with_secion, char *err, size_t errlen);
void psr_tables_free(psr_tables *t);
/* replaces the quadrature SPH table by the reference's data file
 * src/kernel.akline (one value per line, comma-terminated: the array
 * initialiser of src/sph-depth.c:3-6) */
int psr_tables_load_kernel(psr_tables *t, const char *path, char *err, size_t errlen);
/* host table lookups (for tests): ar_get, xs_get, sph_depth */
double psr_tab_ar_get(const psr_tables *t, int col, double logT);
double psr_tab_xs_get(const psr_tables *t, int id, double eng);
double psr_tab_sph_depth(const psr_tables *t, double r_h);

/* LTE recombination spectra (lte_init / lte_fill, src/lte.c:28-72), for
 * the recombination rays: species s (0: HII -> HI, 1: HeII -> HeI, 2:
 * HeIII -> HeII) has PSR_LTE_NFREQ photon energies E_j = E0 + j (16 E0 -
 * E0) / PSR_LTE_NFREQ from its threshold E0 (energy[s][j]), and for each
 * of PSR_LTE_NLOGT rows of log10 T = 3 + i 7/128 the cumulative weights
 * sigma(E_j) E_j^3 exp(-(E_j - E0) / kT) divided by the row's total
 * (cdf[s][i][j]); tbound[k] = 10^(3 + (k + 1) 7/128), k < PSR_LTE_NLOGT - 1:
 * a temperature T takes the row counting the bounds <= T (the reference's
 * (int)((log10 T - 3) / (7/128)) clamped to the rows) */
#define PSR_LTE_NFREQ 1024
#define PSR_LTE_NLOGT 128
#define PSR_LTE_LOGT_MIN 3.0
#define PSR_LTE_LOGT_MAX 10.0
int psr_lte_tables(const psr_tables *t, const psr_units *u, double *energy, double *cdf, double *tbound);

/* ---- snapshots (reader_new, src/reader.c:36-46) ----
 * format: "psphray" (src/reader-psphray.inc; NULL means it) or
 * "massiveblack" / "e5" (src/reader-massiveblack.inc).  header: the
 * 256-byte header record in the psphray field order (struct header_t,
 * src/reader-psphray.inc:1-28, its first 256 bytes); massiveblack's
 * header_t (src/reader-massiveblack.inc:1-23) shares every field up to
 * flag_entropy, which is all its block lengths read.  N[0] gas, N[1] dark
 * matter (massiveblack), N[4] stars, N[5] sources; the flags switch blocks
 * on. */
typedef struct psr_snap_header {
    uint32_t N[6];
    double mass[6];
    double time, redshift;
    int32_t flag_sfr, flag_feedback;
    uint32_t Nparticle_total_low[6];
    int32_t flag_cool, Nfiles;
    double boxsize, OmegaM, OmegaL, h;
    int32_t flag_sft, flag_met;
    uint32_t Nparticle_total_high[6];
    int32_t flag_entropy;
    double OmegaB;
    uint64_t nrays;
    int32_t flag_Hmf, flag_Hemf, flag_Helium, flag_gammaHI, flag_cloudy, flag_eos;
    int32_t unused[4];
} psr_snap_header;
/* a block's length (items), item size and file offset under a header
 * (_length / _itemsize / _offset of each format); psphray blocks: "pos",
 * "id", "mass", "ie", "rho", "ye", "xHI", "sml", "Hmf", "Hemf", "xHeI",
 * "xHeII", "gammaHI", "HIa_cloudy", "eos", "lasthit", "hits", "sft",
 * "ngammas", "spec"; massiveblack: "pos", "vel", "id", "mass", "ie", "rho",
 * "ye", "xHI", "sml", "sfr", "sft", "met", "bhmass", "bhmdot" */
int psr_snapshot_block_info(const char *format, const psr_snap_header *h, const char *block,
                            int64_t *length, int64_t *itemsize, int64_t *offset);
/* the header record (both formats: 256 bytes framed by 256) */
int psr_snapshot_read_header(const char *path, psr_snap_header *h, char *err, size_t errlen);
/* one block into buf (length * itemsize bytes; nothing when the length is
 * 0); the record's size markers are checked (_read) */
int psr_snapshot_read_block(const char *format, const char *path, const psr_snap_header *h,
                            const char *block, void *buf, char *err, size_t errlen);
/* header and every block the header calls for; names[k] / bufs[k] give the
 * blocks' data (a block not given is written as zeros) */
int psr_snapshot_write(const char *format, const char *path, const psr_snap_header *h,
                       const char *const *names, const void *const *bufs, int32_t nblocks, char *err,
                       size_t errlen);

/* ---- epochs (psystem_switch_epoch, src/psystem.c:221-298) ----
 * the next epoch's gas particles, as read from its snapshot */
typedef struct psr_epoch_snap {
    int64_t npar;
    const uint64_t *id;
    const float *pos; /* [npar][3] */
    const float *mass, *sml, *rho, *ie, *xHI;
    const float *ye; /* read by psr_epoch_load only */
} psr_epoch_snap;
typedef struct psr_epoch_stats {
    int64_t matched, added, lost, removed;
    double mean_shift;
} psr_epoch_stats;
/* the first epoch (psystem_read_epoch, src/psystem.c:300-379): s takes the
 * snapshot's gas (npar <= npar_max; the reference sizes npar_max = 1.05 ngas)
 * with x_HI and ye giving the ionization state (x_HeII from the electrons
 * beyond hydrogen's, the rest kept as yeMET); accumulators, lasthit, hits and
 * flags start at zero */
int psr_epoch_load(psr_state *s, uint64_t *id, int64_t npar_max, const psr_epoch_snap *snap,
                   const psr_units *u, int32_t h_only, char *err, size_t errlen);
/* psystem_match_epoch + compress on the host state s (arrays of capacity
 * npar_max, s->npar in use, ids id[]): each snapshot particle takes the
 * nearest current particle (within 500 internal lengths) in its ID bucket
 * (id & (2^28 - 1)), which keeps its ionization state; an unmatched one is
 * appended with the snapshot's x_HI; current particles not matched are
 * removed, the last valid ones filling their places (src/psystem.c:381-504,
 * :109-150); lasthit is rescaled to the new ticks (:266-272).  Then
 * psr_tree_build and psr_state_upload start the epoch on the device. */
int psr_epoch_switch(psr_state *s, uint64_t *id, int64_t npar_max, const psr_epoch_snap *snap,
                     const psr_units *u, int32_t h_only, int32_t periodic, double boxsize,
                     double prev_duration, int64_t prev_nticks, double tick_time,
                     psr_epoch_stats *stats, char *err, size_t errlen);

/* ---- physics set-up ---- */
int psr_set_physics(psr_ctx *ctx, const psr_units *u, const psr_flags *f,
                    const psr_tables *t);

/* ---- epoch: tree + particles ----
 * psr_tree_build replaces rt_switch_epoch (src/raytrace.c:41-151): it
 * rebuilds, on the device, the reference octree (split at the 9th particle,
 * float midpoints, leaf AABB = union of pos +- sml, internal AABB = union of
 * the 8 children) for nfield independent fields laid end to end
 * (field_offsets[nfield+1], boxsizes[nfield]).  It must precede
 * psr_state_upload, whose arrays it reorders. */
int psr_tree_build(psr_ctx *ctx, const float *pos, const float *sml,
                   const int8_t *flag, int64_t npar, int32_t nfield,
                   const int64_t *field_offsets, const double *boxsizes);
/* tree summary: node count and depth; optional export of the node boxes in
 * the device's level order (first_child -1 for leaves) */
int psr_tree_info(psr_ctx *ctx, int64_t *nnodes, int32_t *depth,
                  int64_t *skipped);
int psr_tree_export(psr_ctx *ctx, int64_t cap, float *bot, float *top,
                    int32_t *first_child, int32_t *npar, int32_t *level);
int psr_state_upload(psr_ctx *ctx, const psr_state *s);
int psr_state_download(psr_ctx *ctx, psr_state *s);

/* ---- rays ---- */
/* stores nray rays in the context's device ray pool (replaces the caller's
 * r[] array, src/loop.c:69) */
int psr_rays_upload(psr_ctx *ctx, const psr_ray *rays, int64_t nray);

/* the output blocks the ticks change, computed on the device and gathered
 * into the caller's order (psystem_write_output, src/psystem.c:644-667):
 * x_HI, x_HeI, x_HeII, ye (electrons per hydrogen from H and He,
 * lambda_to_ye) and gammaHI (yGrecHII) as the snapshot's floats, ie,
 * lasthit, hits; npar entries each, NULL skips a block.  32 bytes per
 * particle cross PCIe instead of the whole state. */
int psr_output_gather(psr_ctx *ctx, float *xHI, float *xHeI, float *xHeII, float *ye, float *gammaHI, float *ie,
                      int64_t *lasthit, int32_t *hits);

/* ---- ray emission on the device (emit_rays, src/loop.c:163-373) ----
 * psr_sources_upload: point sources (Source, src/psystem.h:7-24): positions
 * [nsrc][3] (internal length), Ngamma_dot (photons per internal time),
 * field (batched fields; NULL: 0), spectrum index (NULL: 0), lastemit
 * (NULL: 0; src/sources.c:67-78).  Replaces the context's sources. */
int psr_sources_upload(psr_ctx *ctx, int64_t nsrc, const double *pos, const double *ngamma_dot,
                       const int32_t *field, const int32_t *spec, const int64_t *lastemit);
/* spectra (src/spectra.c:50-60): spectrum k has the bins [offsets[k],
 * offsets[k+1]) of photon energies (internal units) and cumulative weights
 * cdf (non-decreasing, last 1); a monochromatic source has one bin */
int psr_spectra_upload(psr_ctx *ctx, int32_t nspec, const int64_t *offsets, const double *energy,
                       const double *cdf);
/* the source rays of one tick, written into the ray pool at [pool_first,
 * pool_first + *nray_out) (the pool grows as needed; a pending tick using
 * that range is finished first).  per_field == 0: nray rays drawn over all
 * sources; else per_field rays for each field with a source.  Streams:
 * Philox4x32-10 keyed (seed, 0x50535248), counter (ray id, tick, block):
 * block 0 word 0 the source, with weight Ngamma_dot * (tick - lastemit) *
 * tick_time; blocks 1.. the isotropic direction (gsl_ran_dir_3d's
 * rejection method, src/loop.c:250); block 0 word 3 the spectrum bin.  Rays
 * in order of source, stably; a source's rays share its weight equally
 * (src/loop.c:305-321); length 2 * boxsize; type -1.  Sources that emitted
 * get lastemit = tick.  Ticks must increase from call to call. */
int psr_emit_rays(psr_ctx *ctx, int64_t tick, double tick_time, int64_t nray, int64_t per_field,
                  uint32_t seed, int64_t pool_first, int64_t *nray_out);
/* recombination rays (emit_rays, src/loop.c:197-232, :273-299, :339-368).
 * Once set with nrec != 0, and when the physics is not on-the-spot,
 * psr_emit_rays appends them after the source rays from tick 2 on
 * (*nray_out counts both).  Candidates: the particles hit by the first
 * nsrc + 1 rays of the last traced tick, nsrc being the source rays
 * psr_emit_rays wrote for it (every ray of an uploaded list): the
 * reference's active set up to x_src_length (src/loop.c:626-637).
 * nrec > 0: nrec rays per species (HII; with helium also HeII, HeIII), each
 * drawn from the candidates (ascending caller index) with weight
 * yGrec_s / x_HI (x_HI = 0 counts as 1e-10), the weights summed in chunks of
 * 1024 candidates, sequentially within a chunk and over the chunk totals;
 * nrec < 0: one ray from each candidate with yGrec_s > threshold
 * (psphray.recombineThreshold, 0.1).  Ray k of species s draws from stream
 * counter (k, tick, (s + 1) << 16 | block): block 0 word 0 the particle,
 * word 3 the photon energy, blocks 1.. the direction.  A ray starts at its
 * particle with yGrec_s N_s (N_s: the particle's hydrogen or helium atoms)
 * shared among the particle's rays of that species, an energy from the LTE
 * spectrum of the particle's temperature (psr_lte_tables; disable_lte: the
 * threshold energy), length 0.2 boxsize, type s, src_or_par its caller
 * index; rays are sorted by (type, particle), a particle's in draw order;
 * then yGrec_s -= Nph / N_s for each ray in order, clamped at 0.  A species
 * whose weights are all zero emits nothing.  Single field only
 * (PSR_ERR_NOTSUP for batched fields). */
int psr_set_recombination(psr_ctx *ctx, int64_t nrec, double threshold, int32_t disable_lte);
/* the sources' lastemit ticks */
int psr_sources_lastemit(psr_ctx *ctx, int64_t *lastemit);
/* a copy of nray rays of the pool from first (the reference keeps r[] on
 * the host, src/loop.c:69; here for checks and ray logs, src/loop.c:80-85) */
int psr_rays_download(psr_ctx *ctx, int64_t first, int64_t nray, psr_ray *rays);

/* parity mode: rt_trace + qsort for each ray of the pool range
 * [ray_first, ray_first + nray) (src/raytrace.c:153-210,
 * src/loop.c:375-386).  Lists are sorted ascending by (d, ipar).  offsets
 * (nray+1) is always filled; when offsets[nray] > cap the call returns
 * PSR_ERR_CAPACITY without writing the lists. */
int psr_trace_lists(psr_ctx *ctx, int64_t ray_first, int64_t nray,
                    int64_t *offsets, int64_t cap, int64_t *ipar, double *d,
                    double *b);

/* one tick over the pool range: trace, deposit, merge, update
 * (src/loop.c:131-148) restricted to the phases in the mask.  tick is the
 * new psys.tick; tick_time the tick length (internal units). */
int psr_tick(psr_ctx *ctx, int64_t ray_first, int64_t nray, int64_t tick,
             double tick_time, int32_t phases, int32_t chem_mode);
/* per-ray outputs of the last deposit: real ray length (r[i].length =
 * x[j-1].d, src/loop.c:594) and photons left (TM) */
int psr_ray_results(psr_ctx *ctx, int64_t nray, double *length, double *tm);
int psr_get_stats(psr_ctx *ctx, psr_stats *st);
int psr_reset_stats(psr_ctx *ctx);
/* phase timers: HIP events on the context stream around trace, sort,
 * deposit and update of every psr_tick while enabled; ms[5] = accumulated
 * trace, sort, deposit, update and whole-tick milliseconds */
int psr_set_timing(psr_ctx *ctx, int enable);
int psr_get_timers(psr_ctx *ctx, double *ms);

/* launch choices of the device path.  No counterpart in the reference (its
 * deposit adds to the particles in place, src/loop.c:486-573); they change
 * how the work is laid out, never the values beyond f32 summation order.
 * PSR_TUNE_DEPOSIT_ACCUM: how a tick's per-particle deposits are summed.
 *   PSR_ACCUM_AUTO   each choice timed (HIP events) on the first asynchronous
 *                    ticks after the tree or physics change; the faster kept
 *   PSR_ACCUM_ATOMIC f32 atomics into the particle arrays
 *   PSR_ACCUM_BINNED records binned by particle range, summed in LDS
 * psr_get_tuning returns the choice in use (PSR_ACCUM_AUTO while timing). */
#define PSR_TUNE_DEPOSIT_ACCUM 1
#define PSR_ACCUM_AUTO 0
#define PSR_ACCUM_ATOMIC 1
#define PSR_ACCUM_BINNED 2
/* PSR_TUNE_TICK_GRAPH: 1 runs each asynchronous tick as one hipGraph
 * launch, recorded once per tick shape (ray count, capacities, phases) and
 * replayed with the tick's values from a device tick record; ticks with
 * phase timers, accumulation trials or hit lists beyond 8192 hits run
 * kernel by kernel.  0: always kernel by kernel.  -1 (default): eight timed
 * trial ticks (two kernel by kernel, two as graph launches, twice; each
 * pair's second tick timed from the first one's end) decide;
 * psr_get_tuning then returns the choice (-1 while undecided).
 * PSR_TUNE_GRAPH_LAUNCHES (read only): ticks run as graph launches so far. */
#define PSR_TUNE_TICK_GRAPH 2
#define PSR_TUNE_GRAPH_LAUNCHES 3
/* PSR_TUNE_TICK_OVERLAP: 1 (default) runs the update of an asynchronous
 * tick launched kernel by kernel (no phase timers) on a second stream, beside
 * the next tick's trace and sort; the next tick's deposit, and every other
 * call, waits for it.  The values are those of consecutive ticks (each tick
 * has its own device record and active-particle array; the next tick's trace
 * overflowing does not stop this tick's update).  0: one stream, the phases
 * of consecutive ticks in order. */
#define PSR_TUNE_TICK_OVERLAP 6
/* PSR_TUNE_DEPOSIT_WALK: how many rays a deposit workgroup carries (their
 * serial transmission walks, src/loop.c:439-589, run one per lane).
 *   PSR_WALK_AUTO    timed together with the accumulation choices (two
 *                    trial deposits each) on the first asynchronous ticks
 *   PSR_WALK_SINGLE  one ray per 3-wavefront workgroup (prep / walk / apply
 *                    pipelined over 128-hit segments in LDS)
 *   PSR_WALK_MULTI   four rays per 4-wavefront workgroup (32-hit segments;
 *                    the four walks in four lanes, two apply wavefronts),
 *                    rays batched by hit count; f32 atomics only
 *   PSR_WALK_MULTI2  the same with two rays (64-hit segments)
 *   PSR_WALK_MULTI8  the same with eight rays (16-hit segments)
 * The values are the same bits either way (deposits within f32 summation
 * order).  The several-ray walks with PSR_ACCUM_BINNED are PSR_ERR_NOTSUP. */
#define PSR_TUNE_DEPOSIT_WALK 4
/* PSR_TUNE_TRACE_THREADS: threads per ray of the traversal, 256 (32 nodes
 * or leaves per step, a 4096-entry stack) or 128 (16, 2048); -1 (default):
 * two trial ticks that trace their rays with both shapes (both launch
 * orders) decide */
#define PSR_TUNE_TRACE_THREADS 5
#define PSR_WALK_AUTO 0
#define PSR_WALK_SINGLE 1
#define PSR_WALK_MULTI 2
#define PSR_WALK_MULTI2 3
#define PSR_WALK_MULTI8 4
/* the autotuner's decided choices kept between runs (no counterpart in the
 * reference): psr_tuning_save writes them to a text file, one line per
 * workload signature (device arch and compute units, particle and field
 * counts, helium, isothermal, on-the-spot, adiabatic): the deposit mode,
 * the tick launch and the trace's threads per ray; replacing that
 * workload's line; psr_tuning_load applies the line matching this
 * context's workload (PSR_ERR_STATE when there is none), so its ticks skip
 * the trials.  Call after psr_tree_build and psr_set_physics, which restart
 * the tuning.  With the environment variable PSR_TUNING_FILE set, every
 * context loads from that file at its first tick and writes each decision
 * back. */
int psr_tuning_save(psr_ctx *ctx, const char *path);
int psr_tuning_load(psr_ctx *ctx, const char *path);
int psr_set_tuning(psr_ctx *ctx, int32_t param, int32_t value);
int psr_get_tuning(psr_ctx *ctx, int32_t param, int32_t *value);

/* step_evolve on the device for n independent steps (src/solve.c:317-323),
 * in place; rtol <= 0 selects the reference tolerances (src/solve.c:151) */
int psr_solve_steps(psr_ctx *ctx, psr_step *steps, int64_t n, double rtol);
/* psr_solve_steps without any device allocation, for callers that cannot
 * allocate (code under HIP graph capture, serving loops, harnesses that
 * forbid hipMalloc).  No counterpart in the reference.  The steps are staged
 * through a fixed per-GPU device array of PSR_NOALLOC_MAX_STEPS steps
 * (4.7 MB, held until the process ends); larger n returns PSR_ERR_CAPACITY
 * without touching the device.  Calls on one GPU share that array: calls
 * that overlap on one GPU (several host threads or contexts) race. */
#define PSR_NOALLOC_MAX_STEPS 32768
int psr_solve_steps_noalloc(psr_ctx *ctx, psr_step *steps, int64_t n, double rtol);

/* ---- learned chemistry (MLP surrogate of step_evolve) ----
 * No counterpart in the reference: a network 22 -> 64 -> 64 -> 64 -> 4
 * (SiLU) replaces step_evolve (src/solve.c:170-224) for one particle-step
 * in PSR_CHEM_SURROGATE ticks, trained on the device (BF16 MFMA forward and
 * backward, FP32 accumulation and Adam) on tuples the exact solver writes.
 * Features (psr_mlp_features): log10 nH, x_HI, x_HeI, x_HeII, log10 T; the
 * absorbed photons and the state they act on: log10(yGdepHI/HeI/HeII +
 * 1e-12), log10(heat + 1e-30), yeMET, logit x_HI/HeI/HeII (clamped to
 * [1e-6, 1 - 1e-6]), eight zero slots; log10 time.  Targets: x_HI, x_HeI, x_HeII and
 * log10 T after the step; the network predicts their change from features
 * 1..4 (logit of the fractions, so they stay inside (0, 1)).  A prediction
 * is used when the features lie in the training domain and the charge /
 * particle-conservation residual x_HeI + x_HeII - 1 (x_HeIII >= 0) is at
 * most the residual threshold, and (balance check, on by default) the
 * trapezoidal ionisation-balance defect of the rate equations over the step
 * (the photo-ionisations deposited against the change plus recombinations);
 * otherwise the exact solver runs for that particle. */
#define PSR_MLP_IN 22
#define PSR_MLP_HID 64
#define PSR_MLP_OUT 4
int psr_mlp_features(const psr_step *s, const psr_units *u, float *x /*22*/);
/* exact chemistry of the following ticks writes one (features, targets) row
 * per solved particle into a device buffer of cap rows (0: off, and frees) */
int psr_mlp_capture(psr_ctx *ctx, int64_t cap);
/* rows captured so far (*n), and the first max_rows of them when x / y */
int psr_mlp_captured(psr_ctx *ctx, int64_t *n, float *x, float *y,
                     int64_t max_rows);
/* train on n (features, targets) rows, or with x = y = NULL on the captured
 * rows (n <= 0: all).  The first call fits the input normalisation and the
 * domain to the data and initialises the parameters from seed; later calls
 * continue from the current parameters.  Adam, cosine-decayed lr; *loss_out
 * = mean summed squared error of the last epoch. */
int psr_mlp_train(psr_ctx *ctx, const float *x, const float *y, int64_t n,
                  int32_t epochs, int32_t batch, float lr, uint64_t seed,
                  float *loss_out);
/* predictions (targets) for n feature rows, before the residual check */
int psr_mlp_forward(psr_ctx *ctx, const float *x, int64_t n, float *y);
/* parameter access: W1[22][64] b1[64] W2[64][64] b2[64] W3[64][64] b3[64]
 * W4[64][4] b4[4] ([in][out] row-major), fp32 master copy */
int64_t psr_mlp_param_count(void);
int psr_mlp_get_params(psr_ctx *ctx, float *p);
int psr_mlp_set_params(psr_ctx *ctx, const float *p);
/* input normalisation (x - mean) / sigma and the domain [lo, hi] per
 * feature (lo / hi may be NULL: unbounded) */
int psr_mlp_set_norm(psr_ctx *ctx, const float *mean, const float *sigma,
                     const float *lo, const float *hi);
int psr_mlp_get_norm(psr_ctx *ctx, float *mean, float *sigma, float *lo,
                     float *hi);
/* gradient of the loss (mean over the n rows of the summed squared error of
 * the predicted changes) at the current parameters, for the gradient check
 * against the FP64 oracle */
int psr_mlp_grad(psr_ctx *ctx, const float *x, const float *y, int64_t n,
                 float *grad, float *loss);
/* residual threshold for the fallback to the exact solver (default 1e-4) */
int psr_mlp_set_residual(psr_ctx *ctx, double tol);
/* also require the ionisation-balance defect within the threshold (fractions)
 * and ten times it relative (ie); default on.  Off, only the conservation
 * residual and the domain guard the prediction: faster, but a network's
 * occasional gross error then goes through (measured up to 1.0 in x_HII). */
int psr_mlp_set_balance_check(psr_ctx *ctx, int32_t on);

#ifdef __cplusplus
}
#endif
#endif
