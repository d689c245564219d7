/*
 * A C caller of the psphray_amd ABI, as a C host (src/main.c + src/loop.c)
 * would drive it: context, units, tables (rates file + src/kernel.akline),
 * physics, tree, particle state, rays, one whole tick, download, stats.
 *
 *   c_driver RATES KERNEL CASE OUT
 *
 * CASE (little-endian, written by tests/test_c_driver.py):
 *   int64 npar, nray, tick; double boxsize, tick_time, hmf;
 *   int32 flags[8] (isothermal adiabatic on_the_spot h_only
 *                   secondary_ionization trace_only no_photon comoving);
 *   double redshift, fake_temperature;
 *   then the particle arrays in psr_state order (pos f32[3n], sml mass rho
 *   ie yeMET f32[n], flag i8[n], lambdaH lambdaHeI lambdaHeII f64[n],
 *   yGdepHI..yGdepHeII yGrecHII..yGrecHeIII heat f32[n], lasthit i64[n],
 *   hits i32[n]) and nray psr_ray records.
 * OUT: the downloaded arrays from lambdaH to hits in the same order, then
 *   psr_stats.
 * Exit status 0 on success; a failing call prints psr_last_error.
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "psphray_amd.h"

#define C_1_CMH 3.0835455558318480e+21
#define C_1_GRAMH 1.9847005219450705e+43
#define C_1_SECH 3.08568025e+16

static int rd(FILE *f, void *p, size_t sz, size_t n) { return fread(p, sz, n, f) == n ? 0 : -1; }
static int wr(FILE *f, const void *p, size_t sz, size_t n) { return fwrite(p, sz, n, f) == n ? 0 : -1; }

#define CHECK(call)                                                                  \
    do {                                                                             \
        int rc_ = (call);                                                            \
        if (rc_) {                                                                   \
            fprintf(stderr, "%s -> %d: %s\n", #call, rc_, ctx ? psr_last_error(ctx) : ""); \
            return 2;                                                                \
        }                                                                            \
    } while (0)

int main(int argc, char **argv) {
    if (argc != 5) {
        fprintf(stderr, "usage: %s RATES KERNEL CASE OUT\n", argv[0]);
        return 1;
    }
    FILE *f = fopen(argv[3], "rb");
    if (!f) return 1;
    int64_t npar, nray, tick;
    double boxsize, tick_time, hmf, redshift, fake_T;
    int32_t fl[8];
    if (rd(f, &npar, 8, 1) || rd(f, &nray, 8, 1) || rd(f, &tick, 8, 1) || rd(f, &boxsize, 8, 1) ||
        rd(f, &tick_time, 8, 1) || rd(f, &hmf, 8, 1) || rd(f, fl, 4, 8) || rd(f, &redshift, 8, 1) ||
        rd(f, &fake_T, 8, 1))
        return 1;
    const size_t n = (size_t)npar;
    psr_state st;
    memset(&st, 0, sizeof st);
    st.npar = npar;
    float **f32[] = {&st.sml, &st.mass, &st.rho, &st.ie, &st.yeMET};
    float **acc[] = {&st.yGdepHI, &st.yGdepHeI, &st.yGdepHeII, &st.yGrecHII,
                     &st.yGrecHeII, &st.yGrecHeIII, &st.heat};
    double **f64[] = {&st.lambdaH, &st.lambdaHeI, &st.lambdaHeII};
    st.pos = malloc(sizeof(float) * 3 * n);
    if (rd(f, st.pos, sizeof(float), 3 * n)) return 1;
    for (int k = 0; k < 5; k++) {
        *f32[k] = malloc(sizeof(float) * n);
        if (rd(f, *f32[k], sizeof(float), n)) return 1;
    }
    st.flag = malloc(n);
    if (rd(f, st.flag, 1, n)) return 1;
    for (int k = 0; k < 3; k++) {
        *f64[k] = malloc(sizeof(double) * n);
        if (rd(f, *f64[k], sizeof(double), n)) return 1;
    }
    for (int k = 0; k < 7; k++) {
        *acc[k] = malloc(sizeof(float) * n);
        if (rd(f, *acc[k], sizeof(float), n)) return 1;
    }
    st.lasthit = malloc(sizeof(int64_t) * n);
    st.hits = malloc(sizeof(int32_t) * n);
    if (rd(f, st.lasthit, sizeof(int64_t), n) || rd(f, st.hits, sizeof(int32_t), n)) return 1;
    psr_ray *rays = malloc(sizeof(psr_ray) * (size_t)(nray > 0 ? nray : 1));
    if (rd(f, rays, sizeof(psr_ray), (size_t)nray)) return 1;
    fclose(f);

    psr_ctx *ctx = NULL;
    psr_units u;
    psr_tables tab;
    memset(&tab, 0, sizeof tab);
    psr_flags flags = {fl[0], fl[1], fl[2], fl[3], fl[4], fl[5], fl[6], fl[7], redshift, fake_T};
    char err[512];
    CHECK(psr_create(&ctx, 0));
    CHECK(psr_units_init(&u, 0.72, hmf, 1.0 - hmf, C_1_CMH, C_1_GRAMH, C_1_SECH));
    if (psr_tables_build(&tab, &u, argv[1], !flags.h_only, err, sizeof err) ||
        psr_tables_load_kernel(&tab, argv[2], err, sizeof err)) {
        fprintf(stderr, "tables: %s\n", err);
        return 2;
    }
    CHECK(psr_set_physics(ctx, &u, &flags, &tab));
    const int64_t off[2] = {0, npar};
    CHECK(psr_tree_build(ctx, st.pos, st.sml, st.flag, npar, 1, off, &boxsize));
    CHECK(psr_state_upload(ctx, &st));
    CHECK(psr_rays_upload(ctx, rays, nray));
    CHECK(psr_reset_stats(ctx));
    CHECK(psr_tick(ctx, 0, nray, tick, tick_time, PSR_PHASE_ALL, PSR_CHEM_EXACT));
    CHECK(psr_state_download(ctx, &st));
    psr_stats stats;
    CHECK(psr_get_stats(ctx, &stats));

    FILE *o = fopen(argv[4], "wb");
    if (!o) return 1;
    for (int k = 0; k < 3; k++)
        if (wr(o, *f64[k], sizeof(double), n)) return 1;
    for (int k = 0; k < 7; k++)
        if (wr(o, *acc[k], sizeof(float), n)) return 1;
    if (wr(o, st.ie, sizeof(float), n) || wr(o, st.lasthit, sizeof(int64_t), n) ||
        wr(o, st.hits, sizeof(int32_t), n) || wr(o, &stats, sizeof stats, 1))
        return 1;
    fclose(o);
    psr_destroy(ctx);
    psr_tables_free(&tab);
    printf("c_driver: %lld particles, %lld rays, %lld intersections\n", (long long)npar,
           (long long)nray, (long long)stats.total_deposit_count);
    return 0;
}
