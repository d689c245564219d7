/*
 * psr_oracle.h -- CPU oracle for the P-SPHRAY photon-packet hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * algorithm (octree build, Pluecker box test, per-ray trace + distance sort,
 * photon deposition, active-set merge and the per-particle chemistry update)
 * used as the checker for the HIP implementation.  Only tests/, the smoke()
 * entry and bench.py's cpu_baseline leg may load it.  The product library
 * (psphray_amd) never links or calls it.
 *
 * Every function cites the reference file:line it restates.
 */
#ifndef PSR_ORACLE_H
#define PSR_ORACLE_H
#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* internal units and constants: src/units.c:66-104, defaults src/config.c:43-51 */
typedef struct {
    double C_H, C_HMF, C_HEMF, C_1_CMH, C_1_GRAMH, C_1_SECH;
    double U_CM, U_GRAM, U_SEC, U_M, U_KM, U_KPC, U_YR, U_MYR, U_KG, U_MSUN,
           U_MPROTON, U_J, U_EV, U_ERG, U_RY_ENG;
    double C_BOLTZMANN, C_PLANCK, C_SPEED_LIGHT, C_H_PER_MASS, C_HE_PER_MASS,
           C_HI_ENERGY, C_HEI_ENERGY, C_HEII_ENERGY,
           U_MPROTON_OVER_C_BOLTZMANN, C_BOLTZMANN_OVER_U_MPROTON;
} or_units;

void or_units_init(or_units *u, double h, double hmf, double hemf,
                   double cmh, double gramh, double sech);

/* run flags: src/config.c:22-42 (CFG_*) */
typedef struct {
    int32_t isothermal, adiabatic, on_the_spot, h_only, secondary_ionization,
            trace_only, no_photon, comoving;
    double redshift;
    double fake_temperature;
} or_flags;

/* one ray packet: struct r_t, src/loop.c:21-35 (plus the field index used by
 * batched runs; 0 for a single field) */
typedef struct {
    float s[3];
    float dir[3];
    double Nph;
    double freq;
    double length;
    int32_t type;
    int32_t field;
    int64_t src_or_par;
} or_ray;

/* particle system SoA view: PSystem, src/psystem.h:32-78 */
typedef struct {
    int64_t npar;
    const float *pos;            /* [npar][3] */
    const float *sml;
    const int8_t *flag;
    float *mass, *rho, *ie, *yeMET;
    double *lambdaH, *lambdaHeI, *lambdaHeII;
    float *yGdepHI, *yGdepHeI, *yGdepHeII;
    float *yGrecHII, *yGrecHeII, *yGrecHeIII;
    float *heat;
    int64_t *lasthit;
    int32_t *hits;
} or_psys;

/* chemistry step: Step, src/psystem.h:84-105 */
typedef struct {
    double lambdaH, lambdaHeI, lambdaHeII, ie, yeMET, yGdepHI, yGdepHeI,
           yGdepHeII, heat, time, nH, T, yGrecHII, yGrecHeII, yGrecHeIII,
           step_remain;
    int32_t refined;
    int32_t nsteps;     /* solver steps taken (diagnostic) */
    int64_t ipar;
} or_step;

/* hit record: Xtype, src/psystem.h:108-112 */
typedef struct {
    int64_t ipar;
    double d;
    double b;
} or_hit;

typedef struct or_ctx or_ctx;

/* context: units + tables (atomic rates, cross sections, kernel, secondary
 * ionization).  rates_path: atomic_rates_*.txt (src/tabfun.c:60-130 format).
 * kernel_table: 4096 doubles of the SPH column kernel (src/sph-depth.c:3-6). */
or_ctx *or_create(const or_units *u, const or_flags *f, const char *rates_path,
                  const double *kernel_table, int32_t kernel_n);
void or_destroy(or_ctx *c);
const char *or_error(const or_ctx *c);

/* lookups */
double or_ar_get(const or_ctx *c, int32_t col, double logT);
double or_xs_get(const or_ctx *c, int32_t id, double eng);
double or_sph_depth(const or_ctx *c, double r_h);
double or_secion_get(or_ctx *c, int32_t id, double E, double log10x);
int32_t or_ar_nrows(const or_ctx *c);
/* copy the scaled rate table: 15 columns x nrows in the fixed column order
 * HIci HIIrcA HIIrccA HIIrcB HIIrccB HeIci HeIIci HeIIrcA HeIIrcB HeIIIrcA
 * HeIIIrcB HIcic HIcec Brem Compton, preceded by the logT column. */
int32_t or_ar_table(const or_ctx *c, double *out /* 16*nrows */, double *min,
                    double *step_inv);
float or_verner_hi(float ryd);
float or_osterbrok_hei(float ryd);
float or_osterbrok_heii(float ryd);

/* src/pluecker-sse.c:28-198 */
int32_t or_pluecker(const float dir[3], const float *dist, const float s2b[3],
                    const float s2t[3]);

/* ---- octree: src/raytrace.c:41-151, 232-316 ---- */
typedef struct {
    int64_t head_par, first_child, parent;
    float bot[3], top[3];
    uint32_t npar;
    int32_t depth;
} or_cell;

typedef struct {
    or_cell *pool;
    int64_t used, size;
    int64_t *next;
    int64_t npar;
    int64_t skipped;
} or_tree;

or_tree *or_tree_build(const float *pos, const float *sml, int64_t npar,
                       double boxsize);
void or_tree_free(or_tree *t);
int64_t or_tree_ncells(const or_tree *t);
/* export: per cell first_child, npar, bot, top, depth (arrays of ncells) */
void or_tree_export(const or_tree *t, int64_t *first_child, int64_t *parent,
                    uint32_t *npar, float *bot, float *top, int32_t *depth);
/* leaf id of every particle (-1 when skipped) */
void or_tree_leaf_of(const or_tree *t, int64_t *leaf);

/* ---- trace: src/raytrace.c:153-230 + sort src/loop.c:375-386 ----
 * Hits sorted ascending by (d, ipar). Two-call protocol: offsets (nray+1)
 * are always filled; returns the total number of hits; hits are written only
 * when total <= cap. */
int64_t or_trace_rays(const or_tree *t, const or_psys *p, const or_ray *rays,
                      int64_t nray, int64_t *offsets, or_hit *hits, int64_t cap,
                      int32_t nthreads);

/* ---- tick phases: src/loop.c:388-776 ---- */
typedef struct {
    int64_t total_deposit_count;     /* sum of full hit-list lengths */
    int64_t saturated_deposit_count;
    int64_t active_count;
    int64_t evolve_count;
    int64_t solver_steps;
    double lost_photon_count_sum;
    double first_ionization_HI, first_ionization_HeI, first_ionization_HeII;
    double secondary_ionization_HI, secondary_ionization_HeI;
    double trace_time, deposit_time, merge_time, update_time;
} or_stats;

/* deposit along sorted hit lists (src/loop.c:388-608).  ray_len_out (nullable)
 * receives the "real ray length" r[i].length = x[j-1].d; ray_tm_out the
 * transmitted photons TM left at the end of each ray. */
int32_t or_deposit(or_ctx *c, or_psys *p, const or_ray *rays, int64_t nray,
                   const int64_t *offsets, const or_hit *hits,
                   double *ray_len_out, double *ray_tm_out, or_stats *st,
                   int32_t nthreads);

/* merge: unique particles in first-hit order (src/loop.c:610-645).  Returns
 * the active count; active must hold offsets[nray] entries. */
int64_t or_merge(const or_psys *p, int64_t nray, const int64_t *offsets,
                 const or_hit *hits, int64_t *active);

/* update: per active particle chemistry (src/loop.c:647-776) */
int32_t or_update(or_ctx *c, or_psys *p, const int64_t *active, int64_t nactive,
                  int64_t tick, double tick_time, or_stats *st, int32_t nthreads);

/* full tick: trace, deposit, merge, update (src/loop.c:131-148) */
int32_t or_tick(or_ctx *c, const or_tree *t, or_psys *p, const or_ray *rays,
                int64_t nray, int64_t tick, double tick_time, or_stats *st,
                int32_t nthreads);

/* chemistry: RHS of src/solve.c:76-141 with the lsoda callback signature
 * (t, x[7], dxdt[7], data) where data points to an or_chem_data. */
typedef struct {
    const or_ctx *c;
    const or_step *step;
} or_chem_data;
int or_chem_rhs(double t, double *x, double *dxdt, void *data);
/* f = or_chem_rhs(x) and jac = df/dx (7 x 7 row-major; may be NULL) by dual
 * numbers, the integrator's Jacobian */
int or_chem_jac(const or_ctx *c, const or_step *step, const double *x, double *f, double *jac);

/* exact one-step solution of the isothermal hydrogen-only on-the-spot rate
 * equation dx/dtau = -g + (1 + m - x)(A - (A + G) x) over tau in [0, 1] */
double or_riccati_tick(double x0, double g, double m, double G, double A);
/* the oracle's implicit FP64 solver (3-stage L-stable Rosenbrock ROS3 with
 * embedded error estimate, same algorithm as the device), integrating
 * normalized time 0..1 and applying the clamps of src/solve.c:208-222; the
 * isothermal hydrogen-only on-the-spot case takes or_riccati_tick instead */
int32_t or_step_evolve(const or_ctx *c, or_step *s, double rtol,
                       const double *atol7);
/* ROS3 with nsub equal steps and no step control (convergence-order tests):
 * final state x7 and the summed embedded error vectors err7 */
int32_t or_ros3_fixed(const or_ctx *c, const or_step *s, int32_t nsub, double *x7, double *err7);
/* chemistry in bulk (for fixtures): n steps in place */
int32_t or_step_evolve_many(const or_ctx *c, or_step *s, int64_t n, double rtol,
                            int32_t nthreads);

/* emit_rays (src/loop.c:163-373) for point sources with counter-based
 * streams (Philox4x32-10, key (seed, 0x50535248), counter (ray id, tick lo,
 * tick hi, block)) in place of the GSL ones.  Groups: every source
 * (per_field == 0, nray rays, ids from 0) or each field's sources
 * (per_field rays, ids from field * per_field).  Writes the rays into out
 * (capacity cap), advances lastemit, returns the ray count (-1: cap). */
int64_t or_emit_rays(int64_t nsrc, const double *pos, const double *ngd, const int32_t *field,
                     const int32_t *spec, int64_t *lastemit, const int64_t *spec_off,
                     const double *spec_eng, const double *spec_cdf, int32_t nfield,
                     const double *boxsizes, int64_t tick, double tick_time, int64_t nray,
                     int64_t per_field, uint32_t seed, or_ray *out, int64_t cap);

/* LTE recombination spectra, lte_fill (src/lte.c:47-72): eng[3][1024]
 * photon energies from each species' threshold E0 to 16 E0; cdf[3][128][1024]
 * the running sums of sigma(E) E^3 exp(-(E - E0)/kT) over a row of
 * log10 T = 3 + i 7/128, divided by the row total; tb[127] the rows' lower
 * bounds in T, 10^(3 + (k+1) 7/128). */
void or_lte_tables(const or_ctx *c, double *eng, double *cdf, double *tb);

/* recombination rays of one tick (emit_rays, src/loop.c:197-232, :273-299,
 * :305-321, :339-368): nothing when on-the-spot, nrec == 0 or tick <= 1.
 * Candidates: the particles the first min(nprev, cut) rays of the previous
 * tick's list hit (the active set up to x_src_length, src/loop.c:626-637),
 * in ascending index.  nrec > 0: per species nrec draws with weights
 * yGrec_s / x_HI (0 -> 1e-10), cumulative weights summed in chunks of 1024
 * (sequentially within a chunk, then over the chunk totals); nrec < 0: every
 * candidate with yGrec_s > threshold.  Draw k of species s uses the stream
 * counter (k, tick lo, tick hi, (s+1) << 16 | block).  Writes the rays
 * sorted by (species, particle), takes the emitted photons off yGrec_s,
 * returns the count (-1: cap). */
int64_t or_emit_recomb(const or_ctx *c, const or_tree *t, or_psys *p, const or_ray *prev, int64_t nprev,
                       int64_t cut, int64_t nrec, double threshold, int32_t disable_lte,
                       const double *lte_eng, const double *lte_cdf, const double *lte_tb, double boxsize,
                       int64_t tick, uint32_t seed, or_ray *out, int64_t cap);

/* units helpers: src/config.h:295-303 */
double or_ieye2T(const or_ctx *c, double ie, double ye);
double or_Tye2ie(const or_ctx *c, double T, double ye);
const or_units *or_get_units(const or_ctx *c);
const or_flags *or_get_flags(const or_ctx *c);
void or_set_flags(or_ctx *c, const or_flags *f);

#ifdef __cplusplus
}
#endif
#endif
