/*
 * psr_oracle.c -- CPU oracle for the P-SPHRAY photon-packet hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  A plain-C restatement of the reference
 * algorithm, used as the checker of the HIP path by tests/, by the smoke()
 * entry and as bench.py's CPU baseline.  The product never links it.
 *
 * Compiled with -ffp-contract=off: the reference geometry relies on
 * separately rounded float products (src/pluecker-sse.c:12-25,
 * src/raytrace.c:175-186).
 */
#define _GNU_SOURCE
#include "psr_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <strings.h>
#include <sys/types.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static double now_s(void) {
#ifdef _OPENMP
    return omp_get_wtime();
#else
    return 0.0;
#endif
}

/* ------------------------------------------------------------------ units */
/* src/units.c:66-104 */
void or_units_init(or_units *u, double h, double hmf, double hemf, double cmh,
                   double gramh, double sech) {
    memset(u, 0, sizeof(*u));
    u->C_H = h;
    u->C_HMF = hmf;
    u->C_HEMF = hemf;
    u->C_1_CMH = cmh;
    u->C_1_GRAMH = gramh;
    u->C_1_SECH = sech;
    u->U_CM = h / cmh;
    u->U_GRAM = h / gramh;
    u->U_SEC = h / sech;
    u->U_M = u->U_CM * 100;
    u->U_KM = u->U_M * 1000;
    u->U_KPC = u->U_M * 3.08568025e19;
    u->U_YR = u->U_SEC * 31556926.0;
    u->U_KG = u->U_GRAM * 1000;
    u->U_MYR = u->U_YR * 1e6;
    u->U_MSUN = u->U_KG * 1.98892e30;
    u->U_MPROTON = u->U_KG * 1.67262158e-27;
    u->C_BOLTZMANN = 1.3806503e-23 * u->U_M * u->U_M * u->U_KG / u->U_SEC / u->U_SEC;
    u->C_PLANCK = 6.626068e-34 * u->U_M * u->U_M * u->U_KG / u->U_SEC;
    u->U_J = u->U_KG * u->U_M * u->U_M / (u->U_SEC * u->U_SEC);
    u->U_EV = 1.602176487e-19 * u->U_J;
    u->U_ERG = 1e-7 * u->U_J;
    u->U_RY_ENG = 13.60569253 * u->U_EV;
    u->C_H_PER_MASS = hmf / u->U_MPROTON;
    u->C_HE_PER_MASS = hemf / (4 * u->U_MPROTON);
    u->C_SPEED_LIGHT = 3e8 * u->U_M / u->U_SEC;
    u->C_HI_ENERGY = 13.60569253 * u->U_EV;
    u->C_HEI_ENERGY = 24.587 * u->U_EV;
    u->C_HEII_ENERGY = 54.416 * u->U_EV;
    u->U_MPROTON_OVER_C_BOLTZMANN = u->U_MPROTON / u->C_BOLTZMANN;
    u->C_BOLTZMANN_OVER_U_MPROTON = 1.0 / (u->U_MPROTON / u->C_BOLTZMANN);
}

/* --------------------------------------------------------------- tabfun */
/* TabFun, src/tabfun.h:1-10 */
typedef struct {
    int nrows, ncols, cap;
    double min, max, step, step_inv;
    double **data;
    char **headers;
} or_tab;

/* src/tabfun.c:38-58 */
static double tab_get(const or_tab *t, int id, double value) {
    ssize_t index = (ssize_t)((value - t->min) * t->step_inv);
    if (index < 0) return t->data[id][0];
    if (index >= t->nrows - 1) return t->data[id][t->nrows - 1];
    double left = t->data[0][index];
    double right = t->data[0][index + 1];
    double leftwt = (right - value);
    double rightwt = (value - left);
    return (leftwt * t->data[id][index] + rightwt * t->data[id][index + 1]) * t->step_inv;
}

/* src/tabfun.c:60-130: fixed 14-character Fortran columns */
static int tab_load(or_tab *t, const char *path, char *err) {
    FILE *fp = fopen(path, "r");
    if (!fp) { snprintf(err, 256, "cannot open %s", path); return -1; }
    char *line = NULL;
    size_t size = 0;
    ssize_t len;
    int stage = 0, irow = 0;
    memset(t, 0, sizeof(*t));
    while ((len = getline(&line, &size, fp)) >= 0) {
        if (line[0] == '#') continue;
        if (stage == 0) {
            if (4 != sscanf(line, "%lf %lf %d %lf", &t->min, &t->max, &t->nrows, &t->step)) {
                snprintf(err, 256, "bad header in %s", path);
                fclose(fp); free(line); return -1;
            }
            t->step_inv = 1.0 / t->step;
            stage = 1;
        } else if (stage == 1) {
            t->ncols = (int)(len / 14);
            t->cap = t->ncols + 4;
            t->data = calloc(t->cap, sizeof(double *));
            t->headers = calloc(t->cap, sizeof(char *));
            for (int i = 0; i < t->ncols; i++) {
                char *p = line + i * 14;
                char *q = p + 13;
                while (q >= p && *q == ' ') { *q = 0; q--; }
                while (*p == ' ') p++;
                t->headers[i] = strdup(p);
                t->data[i] = calloc(t->nrows, sizeof(double));
            }
            stage = 2;
        } else if (stage == 2) {
            stage = 3;
            irow = 0;
        } else if (stage == 3) {
            for (int i = 0; i < t->ncols; i++) {
                char *p = line + i * 14;
                if (p > line + len) {
                    snprintf(err, 256, "short line in %s", path);
                    fclose(fp); free(line); return -1;
                }
                p[13] = 0;
                t->data[i][irow] = atof(p);
            }
            irow++;
            if (irow == t->nrows) break;
        }
    }
    free(line);
    fclose(fp);
    if (irow != t->nrows) { snprintf(err, 256, "table %s truncated", path); return -1; }
    return 0;
}

static void tab_free(or_tab *t) {
    int n = t->cap;
    if (t->data) for (int i = 0; i < n; i++) free(t->data[i]);
    if (t->headers) for (int i = 0; i < n; i++) free(t->headers[i]);
    free(t->data);
    free(t->headers);
}

/* src/tabfun.c:12-36: find a column by name and scale it, else synthesize it
 * from an analytic form evaluated on column 0 */
static int tab_setup_col(or_tab *t, const char *col, double (*func)(double), double unit,
                         char *err) {
    for (int i = 0; i < t->ncols; i++) {
        if (t->headers[i] && !strcasecmp(col, t->headers[i])) {
            for (int j = 0; j < t->nrows; j++) t->data[i][j] *= unit;
            return i;
        }
    }
    if (!func) { snprintf(err, 256, "column %s unknown", col); return -1; }
    if (t->ncols >= t->cap) { snprintf(err, 256, "too many columns"); return -1; }
    int i = t->ncols++;
    t->headers[i] = strdup(col);
    t->data[i] = calloc(t->nrows, sizeof(double));
    for (int j = 0; j < t->nrows; j++) t->data[i][j] = func(t->data[0][j]) * unit;
    return i;
}

/* src/atomic-rate.c:245-251: note that tabfun_setup_col passes column 0
 * (log T) to the analytic form */
static double brem_cen_1992(double T) { return 1.42e-27 * 1.5 * sqrt(T); }
static double compton_haiman_1996(double T) { (void)T; return 0.0; }

/* ------------------------------------------------- cross sections (F90) */
/* src/verner.F90:5-32, single precision with a double exponent */
float or_verner_hi(float Ry) {
    const float E0 = (float)4.298e-1, sig0 = (float)5.475e4, ya = (float)3.288e1,
                P = (float)2.963;
    if (Ry < 0.5f) return 0.0f;
    float eV = (float)((double)Ry * 13.6);
    float x = eV / E0;
    float y = x;
    const float xm1 = x - 1.0f;
    float a = sig0 * (xm1 * xm1);
    double b = pow((double)y, 0.5 * (double)P - 5.5);
    float c = powf(1.0f + sqrtf(y / ya), -P);
    float sigma = (float)((double)a * b * (double)c);
    sigma = (float)((double)sigma * 1.0e-18);
    if (sigma < 0) sigma = 0.0f;
    return sigma;
}

/* src/verner.F90:36-50 */
float or_osterbrok_hei(float Ry) {
    const float nu = 24.587f;
    if (2.0f * Ry * 13.6f < nu) return 0.0f;
    float sf = Ry * 13.6f / nu;
    return 7.2e-18f * (1.66f * powf(sf, -2.05f) + 0.66f * powf(sf, -3.05f));
}

/* src/verner.F90:54-64 */
float or_osterbrok_heii(float Ry) {
    const float nu = 54.416f;
    if (2.0f * Ry * 13.6f < nu) return 0.0f;
    float t = Ry * 13.6f / nu;
    return 1.58e-18f * (1.0f / (t * t * t));
}

/* --------------------------------------------------------------- context */
enum {
    AR_HI_CI = 0, AR_HII_RC_A, AR_HII_RCC_A, AR_HII_RC_B, AR_HII_RCC_B, AR_HEI_CI,
    AR_HEII_CI, AR_HEII_RC_A, AR_HEII_RC_B, AR_HEIII_RC_A, AR_HEIII_RC_B, AR_HI_CIC,
    AR_HI_CEC, AR_E_BREMC, AR_E_COMPC, AR_NCOL
};

struct or_ctx {
    or_units u;
    or_flags f;
    or_tab ar;
    int arcol[AR_NCOL];
    or_tab xs;
    int xscol[3];
    double xs_cut[3];
    double *kernel;
    int kernel_n;
    /* secondary ionization TabFun2 (src/secion.c:75-99), built lazily */
    int si_ready;
    int si_n1, si_n2;
    double si_min1, si_min2, si_step1_inv, si_step2_inv;
    double *si_x, *si_y, *si_data[3];
    char err[256];
};

const char *or_error(const or_ctx *c) { return c->err; }
const or_units *or_get_units(const or_ctx *c) { return &c->u; }
const or_flags *or_get_flags(const or_ctx *c) { return &c->f; }
void or_set_flags(or_ctx *c, const or_flags *f) { c->f = *f; }

static const or_units *g_u_for_xs; /* tab_setup_col takes a plain function */
static double xs_verner(double eng) {
    float ryd = (float)(eng / g_u_for_xs->C_HI_ENERGY);
    return or_verner_hi(ryd);
}
static double xs_hei(double eng) {
    float ryd = (float)(eng / g_u_for_xs->C_HI_ENERGY);
    return or_osterbrok_hei(ryd);
}
static double xs_heii(double eng) {
    float ryd = (float)(eng / g_u_for_xs->C_HI_ENERGY);
    return or_osterbrok_heii(ryd);
}

or_ctx *or_create(const or_units *u, const or_flags *f, const char *rates_path,
                  const double *kernel_table, int32_t kernel_n) {
    or_ctx *c = calloc(1, sizeof(or_ctx));
    c->u = *u;
    c->f = *f;
    char *err = c->err;
    /* src/atomic-rate.c:253-274 */
    if (tab_load(&c->ar, rates_path, err)) return c;
    const double cm3s = u->U_CM * u->U_CM * u->U_CM / u->U_SEC;
    const double ecm3s = u->U_ERG * u->U_CM * u->U_CM * u->U_CM / u->U_SEC;
    static const char *names[AR_NCOL] = {"HIci", "HIIrcA", "HIIrccA", "HIIrcB", "HIIrccB",
                                         "HeIci", "HeIIci", "HeIIrcA", "HeIIrcB", "HeIIIrcA",
                                         "HeIIIrcB", "HIcic", "HIcec", "Brem", "Compton"};
    const double units[AR_NCOL] = {cm3s, cm3s, ecm3s, cm3s, ecm3s, cm3s, cm3s, cm3s,
                                   cm3s, cm3s, cm3s, ecm3s, ecm3s, ecm3s,
                                   u->U_ERG / u->U_SEC};
    for (int k = 0; k < AR_NCOL; k++) {
        double (*fn)(double) = NULL;
        if (k == AR_E_BREMC) fn = brem_cen_1992;
        if (k == AR_E_COMPC) fn = compton_haiman_1996;
        c->arcol[k] = tab_setup_col(&c->ar, names[k], fn, units[k], err);
        if (c->arcol[k] < 0) return c;
    }
    /* src/atomic-rate.c:280-311 (analytic table, no xs file) */
    or_tab *x = &c->xs;
    x->ncols = 1;
    x->nrows = 16384;
    x->min = 1 * u->C_HI_ENERGY;
    x->max = 100 * u->C_HI_ENERGY;
    x->step = (x->max - x->min) / (x->nrows - 1);
    x->step_inv = 1.0 / x->step;
    x->cap = 8;
    x->data = calloc(8, sizeof(double *));
    x->headers = calloc(8, sizeof(char *));
    x->headers[0] = strdup("eng");
    x->data[0] = malloc(sizeof(double) * x->nrows);
    for (int i = 0; i < x->nrows; i++) x->data[0][i] = x->min + x->step * i;
    g_u_for_xs = &c->u;
    c->xscol[0] = tab_setup_col(x, "HI", xs_verner, u->U_CM * u->U_CM, err);
    c->xscol[1] = tab_setup_col(x, "HeI", xs_hei, u->U_CM * u->U_CM, err);
    c->xscol[2] = tab_setup_col(x, "HeII", xs_heii, u->U_CM * u->U_CM, err);
    c->xs_cut[0] = 13.6 * u->U_EV;
    c->xs_cut[1] = 24.5 * u->U_EV;
    c->xs_cut[2] = 54.4 * u->U_EV;
    c->kernel_n = kernel_n;
    c->kernel = malloc(sizeof(double) * kernel_n);
    memcpy(c->kernel, kernel_table, sizeof(double) * kernel_n);
    c->err[0] = 0;
    return c;
}

void or_destroy(or_ctx *c) {
    if (!c) return;
    tab_free(&c->ar);
    tab_free(&c->xs);
    free(c->kernel);
    free(c->si_x);
    free(c->si_y);
    for (int k = 0; k < 3; k++) free(c->si_data[k]);
    free(c);
}

double or_ar_get(const or_ctx *c, int32_t col, double logT) {
    return tab_get(&c->ar, c->arcol[col], logT);
}

/* src/atomic-rate.c:313-318 */
double or_xs_get(const or_ctx *c, int32_t id, double eng) {
    if (eng < c->xs_cut[id]) return 0.0;
    return tab_get(&c->xs, c->xscol[id], eng);
}

/* src/sph-depth.c:224-231 */
double or_sph_depth(const or_ctx *c, double r_h) {
    const int N = c->kernel_n;
    int n = (int)(N * r_h);
    if (n < N && n >= 0) return c->kernel[n];
    return 0.0;
}

int32_t or_ar_nrows(const or_ctx *c) { return c->ar.nrows; }

int32_t or_ar_table(const or_ctx *c, double *out, double *min, double *step_inv) {
    int n = c->ar.nrows;
    memcpy(out, c->ar.data[0], sizeof(double) * n);
    for (int k = 0; k < AR_NCOL; k++)
        memcpy(out + (size_t)(k + 1) * n, c->ar.data[c->arcol[k]], sizeof(double) * n);
    *min = c->ar.min;
    *step_inv = c->ar.step_inv;
    return n;
}

/* ------------------------------------------------ secondary ionization */
/* src/secion.c:21-73 (Ricotti, Gnedin & Shull 2002 fits, as tabulated) */
static const double SI_a[3][2] = {{NAN, 0.2}, {NAN, 0.2}, {NAN, 0.4}};
static const double SI_b[3][2] = {{0.4092, 0.38}, {0.4614, 0.38}, {0.2663, 0.34}};
static const double SI_c[3][2] = {{1.7592, 2.0}, {1.6660, 2.0}, {1.3163, 2.0}};
static const double SI_C[3][2] = {{0.3908, 0.6941}, {0.0554, 0.0984}, {1.0, 3.9811}};

static double si_phi(const or_units *u, int j, double E, double x) {
    double gate = 28 * u->U_EV;
    int f1 = (E >= (gate - 0.1));
    double f2 = f1 * pow(gate / E, 0.4);
    double y1 = SI_C[j][0] * pow(1. - pow(x, SI_b[j][0]), SI_c[j][0]);
    double y2 = SI_C[j][1] * pow(x, SI_a[j][1]) * pow(1 - pow(x, SI_b[j][1]), SI_c[j][1]);
    return (y1 * f1 - y2 * f2);
}
static double si_eh(const or_units *u, double E, double x) {
    double gate = 11 * u->U_EV;
    int f1 = (E >= gate - 0.1);
    double f2 = f1 * pow(gate / E, 0.7);
    double y1 = SI_C[2][0] * pow(1 - pow(x, SI_b[2][0]), SI_c[2][0]);
    double y2 = SI_C[2][1] * pow(x, SI_a[2][1]) * pow(1 - pow(x, SI_b[2][1]), SI_c[2][1]);
    return (1. - y1 * f1 + y2 * f2);
}

/* src/secion.c:75-99 */
static void si_init(or_ctx *c) {
    if (c->si_ready) return;
    const or_units *u = &c->u;
    c->si_n1 = 4096;
    c->si_n2 = 1024;
    c->si_min1 = 10 * u->U_EV;
    double max1 = 200 * u->U_EV;
    c->si_min2 = -8;
    double max2 = 0;
    double step1 = (max1 - c->si_min1) / c->si_n1;
    c->si_step1_inv = 1.0 / step1;
    double step2 = (max2 - c->si_min2) / c->si_n2;
    c->si_step2_inv = 1.0 / step2;
    c->si_x = malloc(sizeof(double) * c->si_n1);
    c->si_y = malloc(sizeof(double) * c->si_n2);
    for (int i = 0; i < c->si_n1; i++) c->si_x[i] = c->si_min1 + i * step1;
    for (int i = 0; i < c->si_n2; i++) c->si_y[i] = c->si_min2 + i * step2;
    for (int k = 0; k < 3; k++) c->si_data[k] = malloc(sizeof(double) * c->si_n1 * c->si_n2);
    for (int i = 0; i < c->si_n1; i++)
        for (int j = 0; j < c->si_n2; j++) {
            double x = pow(10, c->si_y[j]);
            size_t o = (size_t)i * c->si_n2 + j;
            c->si_data[0][o] = si_phi(u, 0, c->si_x[i], x);
            c->si_data[1][o] = si_phi(u, 1, c->si_x[i], x);
            c->si_data[2][o] = si_eh(u, c->si_x[i], x);
        }
    c->si_ready = 1;
}

/* src/tabfun.c:173-223 */
double or_secion_get(or_ctx *c, int32_t id, double x, double y) {
    si_init(c);
    const double *d = c->si_data[id];
    const int n1 = c->si_n1, n2 = c->si_n2;
    int xind = (int)((x - c->si_min1) * c->si_step1_inv);
    int yind = (int)((y - c->si_min2) * c->si_step2_inv);
    int linoffset = -1;
    if (xind < 0) linoffset = 0;
    if (xind >= n1 - 1) linoffset = (n1 - 1) * n2;
    if (linoffset != -1) {
        if (yind < 0) return d[linoffset + 0];
        if (yind >= n2 - 1) return d[linoffset + n2 - 1];
        double left = c->si_y[yind], right = c->si_y[yind + 1];
        double leftwt = (right - y), rightwt = (y - left);
        return (leftwt * d[linoffset + yind] + rightwt * d[linoffset + yind + 1]) * c->si_step2_inv;
    }
    int lin = -1;
    if (yind < 0) lin = 0;
    if (yind >= n2 - 1) lin = n2 - 1;
    if (lin != -1) {
        double left = c->si_x[xind], right = c->si_x[xind + 1];
        double leftwt = (right - x), rightwt = (x - left);
        return (leftwt * d[(size_t)xind * n2 + lin] + rightwt * d[(size_t)(xind + 1) * n2 + lin]) *
               c->si_step1_inv;
    }
    double x1 = c->si_x[xind], x2 = c->si_x[xind + 1];
    double y1 = c->si_y[yind], y2 = c->si_y[yind + 1];
    double fac = c->si_step1_inv * c->si_step2_inv;
    double f11 = d[(size_t)xind * n2 + yind];
    double f12 = d[(size_t)xind * n2 + yind + 1];
    double f21 = d[(size_t)(xind + 1) * n2 + yind];
    double f22 = d[(size_t)(xind + 1) * n2 + yind + 1];
    return fac * (f11 * (x2 - x) * (y2 - y) + f21 * (x - x1) * (y2 - y) +
                  f12 * (x2 - x) * (y - y1) + f22 * (x - x1) * (y - y1));
}

/* ------------------------------------------------------------- pluecker */
/* src/pluecker-sse.c:28-198.  The eight direction classes of the reference
 * share one rule per axis pair (i, j): with s_i = dir[i] > 0,
 *   A = s_i ? t_j : b_j,  B = s_j ? b_i : t_i,
 *   C = s_i ? b_j : t_j,  D = s_j ? t_i : b_i,
 * reject when dir[i]*A < dir[j]*B or dir[i]*C > dir[j]*D (float products).
 * The start/end slab rejections pick top or bottom by the same sign. */
int32_t or_pluecker(const float dir[3], const float *dist, const float s2b[3],
                    const float s2t[3]) {
    float e2t[3], e2b[3];
    int sg[3];
    for (int i = 0; i < 3; i++) {
        float m = *dist * dir[i];
        e2t[i] = s2t[i] - m;
        e2b[i] = s2b[i] - m;
        sg[i] = dir[i] > 0;
    }
    for (int i = 0; i < 3; i++) {
        if (sg[i] ? (s2t[i] < 0) : (s2b[i] > 0)) return 0;
    }
    for (int i = 0; i < 3; i++) {
        if (sg[i] ? (e2b[i] > 0) : (e2t[i] < 0)) return 0;
    }
    static const int P[3][2] = {{0, 1}, {0, 2}, {1, 2}};
    for (int k = 0; k < 3; k++) {
        const int i = P[k][0], j = P[k][1];
        const float A = sg[i] ? s2t[j] : s2b[j];
        const float B = sg[j] ? s2b[i] : s2t[i];
        const float C = sg[i] ? s2b[j] : s2t[j];
        const float D = sg[j] ? s2t[i] : s2b[i];
        volatile float r0 = dir[i] * A, r1 = dir[j] * B, r2 = dir[i] * C, r3 = dir[j] * D;
        if (r0 < r1 || r2 > r3) return 0;
    }
    return 1;
}

/* ---------------------------------------------------------------- octree */
#define OR_PPC 8 /* CFG_OCTTREE_PPC, src/config.c:32 */

static int cell_inside(const or_tree *t, const float *pos, int64_t icell) {
    const or_cell *c = &t->pool[icell];
    for (int d = 0; d < 3; d++)
        if (pos[d] >= c->top[d] || pos[d] < c->bot[d]) return 0;
    return 1;
}

static void cell_add(or_tree *t, int64_t ipar, int64_t icell) {
    or_cell *c = &t->pool[icell];
    t->next[ipar] = c->head_par;
    c->head_par = ipar;
    c->npar++;
}

static int64_t cell_find(const or_tree *t, const float *pos, int64_t icell) {
    for (;;) {
        const or_cell *c = &t->pool[icell];
        if (c->first_child == -1) return icell;
        int64_t nxt = -1;
        for (int i = 0; i < 8; i++)
            if (cell_inside(t, pos, c->first_child + i)) { nxt = c->first_child + i; break; }
        if (nxt < 0) return -1;
        icell = nxt;
    }
}

/* src/raytrace.c:268-305 */
static void cell_split(or_tree *t, int64_t icell, const float *pos) {
    if (t->used + 8 > t->size) {
        t->size = t->size * 2 + 64;
        t->pool = realloc(t->pool, sizeof(or_cell) * t->size);
    }
    or_cell *cell = &t->pool[icell];
    cell->first_child = t->used;
    t->used += 8;
    or_cell *fc = &t->pool[cell->first_child];
    float center[3];
    for (int d = 0; d < 3; d++) center[d] = (float)(0.5 * (double)(cell->top[d] + cell->bot[d]));
    for (int i = 0; i < 8; i++) {
        fc[i].parent = icell;
        for (int d = 0; d < 3; d++) {
            fc[i].bot[d] = ((1 << d) & i) ? center[d] : cell->bot[d];
            fc[i].top[d] = ((1 << d) & i) ? cell->top[d] : center[d];
        }
        fc[i].head_par = -1;
        fc[i].first_child = -1;
        fc[i].npar = 0;
        fc[i].depth = cell->depth + 1;
    }
    cell->npar = 0;
    int64_t nextpar;
    for (int64_t ipar = cell->head_par; ipar != -1; ipar = nextpar) {
        nextpar = t->next[ipar];
        for (int i = 0; i < 8; i++)
            if (cell_inside(t, pos + 3 * ipar, i + cell->first_child)) cell_add(t, ipar, i + cell->first_child);
    }
    cell->head_par = -1;
}

/* src/raytrace.c:41-151 */
or_tree *or_tree_build(const float *pos, const float *sml, int64_t npar, double boxsize) {
    or_tree *t = calloc(1, sizeof(or_tree));
    t->npar = npar;
    t->size = npar / OR_PPC + 64;
    t->pool = malloc(sizeof(or_cell) * t->size);
    t->next = malloc(sizeof(int64_t) * (npar > 0 ? npar : 1));
    for (int64_t i = 0; i < npar; i++) t->next[i] = -1;
    or_cell *root = &t->pool[0];
    root->parent = -1;
    for (int d = 0; d < 3; d++) {
        root->bot[d] = 0;
        root->top[d] = (float)boxsize;
    }
    root->first_child = -1;
    root->head_par = -1;
    root->npar = 0;
    root->depth = 0;
    t->used = 1;
    int64_t icell = 0;
    for (int64_t ipar = 0; ipar < npar; ipar++) {
        const float *p = pos + 3 * ipar;
        while (!cell_inside(t, p, icell)) {
            icell = t->pool[icell].parent;
            if (icell == -1) break;
        }
        if (icell == -1) {
            t->skipped++;
            icell = 0;
            continue;
        }
        icell = cell_find(t, p, icell);
        while (t->pool[icell].npar >= OR_PPC) {
            if (t->pool[icell].depth > 40) break; /* coincident particles */
            cell_split(t, icell, pos);
            icell = cell_find(t, p, icell);
        }
        cell_add(t, ipar, icell);
    }
    /* bottom-up AABB refit, src/raytrace.c:100-150: children always have a
     * larger index than their parent, so a reverse sweep is a valid order */
    for (int64_t ic = t->used - 1; ic >= 0; ic--) {
        or_cell *c = &t->pool[ic];
        int first = 1;
        if (c->first_child == -1) {
            for (int64_t ip = c->head_par; ip != -1; ip = t->next[ip]) {
                const float *p = pos + 3 * ip;
                const float h = sml[ip];
                for (int d = 0; d < 3; d++) {
                    if (first || p[d] - h < c->bot[d]) c->bot[d] = p[d] - h;
                    if (first || p[d] + h > c->top[d]) c->top[d] = p[d] + h;
                }
                first = 0;
            }
        } else {
            for (int i = 0; i < 8; i++) {
                const or_cell *ch = &t->pool[c->first_child + i];
                for (int d = 0; d < 3; d++) {
                    if (first || ch->bot[d] < c->bot[d]) c->bot[d] = ch->bot[d];
                    if (first || ch->top[d] > c->top[d]) c->top[d] = ch->top[d];
                }
                first = 0;
            }
        }
    }
    return t;
}

void or_tree_free(or_tree *t) {
    if (!t) return;
    free(t->pool);
    free(t->next);
    free(t);
}

int64_t or_tree_ncells(const or_tree *t) { return t->used; }

void or_tree_export(const or_tree *t, int64_t *first_child, int64_t *parent, uint32_t *npar,
                    float *bot, float *top, int32_t *depth) {
    for (int64_t i = 0; i < t->used; i++) {
        const or_cell *c = &t->pool[i];
        if (first_child) first_child[i] = c->first_child;
        if (parent) parent[i] = c->parent;
        if (npar) npar[i] = c->npar;
        if (depth) depth[i] = c->depth;
        for (int d = 0; d < 3; d++) {
            if (bot) bot[3 * i + d] = c->bot[d];
            if (top) top[3 * i + d] = c->top[d];
        }
    }
}

void or_tree_leaf_of(const or_tree *t, int64_t *leaf) {
    for (int64_t i = 0; i < t->npar; i++) leaf[i] = -1;
    for (int64_t ic = 0; ic < t->used; ic++) {
        const or_cell *c = &t->pool[ic];
        if (c->first_child != -1) continue;
        for (int64_t ip = c->head_par; ip != -1; ip = t->next[ip]) leaf[ip] = ic;
    }
}

/* ----------------------------------------------------------------- trace */
static int cell_hit(const or_tree *t, const float s[3], const float dir[3], float dist,
                    int64_t icell) {
    const or_cell *c = &t->pool[icell];
    float s2b[3], s2t[3];
    for (int d = 0; d < 3; d++) {
        s2b[d] = c->bot[d] - s[d];
        s2t[d] = c->top[d] - s[d];
    }
    return or_pluecker(dir, &dist, s2b, s2t);
}

static int hit_cmp(const void *a, const void *b) {
    const or_hit *x = a, *y = b;
    if (x->d < y->d) return -1;
    if (x->d > y->d) return 1;
    return (x->ipar > y->ipar) - (x->ipar < y->ipar);
}

/* src/raytrace.c:153-210 */
static int64_t trace_one(const or_tree *t, const float *pos, const float *sml,
                         const int8_t *flag, const float s[3], const float dir[3], float dist,
                         or_hit **x, int64_t *size) {
    int64_t length = 0;
    if (t->used == 0) return 0;
    int64_t icell = 0;
    while (icell != -1) {
        if (cell_hit(t, s, dir, dist, icell)) {
            if (t->pool[icell].first_child != -1) {
                icell = t->pool[icell].first_child;
                continue;
            }
            for (int64_t ip = t->pool[icell].head_par; ip != -1; ip = t->next[ip]) {
                if (flag && (flag[ip] & 1)) continue;
                const float *p = pos + 3 * ip;
                const float h = sml[ip];
                double dd2 = 0.0, proj = 0.0;
                for (int d = 0; d < 3; d++) {
                    const float dd = p[d] - s[d];
                    const float pr = dd * dir[d];
                    const float sq = dd * dd;
                    proj += (double)pr;
                    dd2 += (double)sq;
                }
                const double dst = sqrt(dd2);
                const double r2 = (dst - proj) * (dst + proj);
                const float hh = h * h;
                if ((double)hh < r2) continue;
                if (length == *size) {
                    *size = *size * 2 + 16;
                    *x = realloc(*x, sizeof(or_hit) * *size);
                }
                (*x)[length].ipar = ip;
                (*x)[length].d = dst;
                (*x)[length].b = sqrt(fabs(r2));
                length++;
            }
        }
        int64_t next = -1;
        while (icell != 0) {
            const int64_t par = t->pool[icell].parent;
            if (par != -1 && icell - t->pool[par].first_child != 7) { next = icell + 1; break; }
            icell = par;
        }
        icell = next;
    }
    return length;
}

/* all rays: per-ray lists sorted by (d, ipar), concatenated (malloc'd) */
static or_hit *trace_all(const or_tree *t, const or_psys *p, const or_ray *rays, int64_t nray,
                         int64_t *offsets, int32_t nthreads) {
    or_hit **lists = calloc(nray > 0 ? nray : 1, sizeof(or_hit *));
    int64_t *lens = calloc(nray > 0 ? nray : 1, sizeof(int64_t));
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = 1;
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
#endif
    for (int64_t i = 0; i < nray; i++) {
        int64_t size = 0;
        or_hit *x = NULL;
        int64_t n = trace_one(t, p->pos, p->sml, p->flag, rays[i].s, rays[i].dir,
                              (float)rays[i].length, &x, &size);
        qsort(x, n, sizeof(or_hit), hit_cmp);
        lists[i] = x;
        lens[i] = n;
    }
    offsets[0] = 0;
    for (int64_t i = 0; i < nray; i++) offsets[i + 1] = offsets[i] + lens[i];
    const int64_t total = offsets[nray];
    or_hit *hits = malloc(sizeof(or_hit) * (total > 0 ? total : 1));
    for (int64_t i = 0; i < nray; i++)
        if (lens[i]) memcpy(hits + offsets[i], lists[i], sizeof(or_hit) * lens[i]);
    for (int64_t i = 0; i < nray; i++) free(lists[i]);
    free(lists);
    free(lens);
    return hits;
}

int64_t or_trace_rays(const or_tree *t, const or_psys *p, const or_ray *rays, int64_t nray,
                      int64_t *offsets, or_hit *hits, int64_t cap, int32_t nthreads) {
    or_hit *all = trace_all(t, p, rays, nray, offsets, nthreads);
    const int64_t total = offsets[nray];
    if (hits && total <= cap && total > 0) memcpy(hits, all, sizeof(or_hit) * total);
    free(all);
    return total;
}

/* ---------------------------------------------------------------- deposit */
static inline double psys_ye(const or_ctx *c, const or_psys *p, int64_t i) {
    /* lambda_to_ye, src/psystem.h:134-138 */
    const double lH = p->lambdaH[i], lHeI = p->lambdaHeI[i], lHeII = p->lambdaHeII[i];
    return (1.0 - lH) + c->u.C_HEMF / c->u.C_HMF * 0.25 * (lHeII + 2 * (1.0 - lHeII - lHeI));
}

static inline void add_f32(float *x, double v, int atomic) {
    if (atomic) {
#ifdef _OPENMP
#pragma omp atomic
#endif
        *x += v;
    } else {
        *x = (float)((double)*x + v);
    }
}
static inline void add_i32(int32_t *x, int atomic) {
    if (atomic) {
#ifdef _OPENMP
#pragma omp atomic
#endif
        (*x)++;
    } else {
        (*x)++;
    }
}

static double fdim_(double a, double b) { return a > b ? a - b : 0.0; }

/* src/loop.c:388-608 */
int32_t or_deposit(or_ctx *c, or_psys *p, const or_ray *rays, int64_t nray,
                   const int64_t *offsets, const or_hit *hits, double *ray_len_out,
                   double *ray_tm_out, or_stats *st, int32_t nthreads) {
    const or_units *u = &c->u;
    const or_flags *f = &c->f;
    const double scaling_fac = f->comoving ? 1 / (f->redshift + 1) : 1.0;
    const double scaling_fac2_inv = 1.0 / (scaling_fac * scaling_fac);
    if (f->secondary_ionization) si_init(c);
    int64_t saturated = 0, total = 0;
    double fiHI = 0, fiHeI = 0, fiHeII = 0, siHI = 0, siHeI = 0, lost = 0;
    const int atomic = nthreads > 1;
    double t0 = now_s();
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = 1;
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads) \
    reduction(+ : saturated, total, fiHI, fiHeI, fiHeII, siHI, siHeI, lost)
#endif
    for (int64_t i = 0; i < nray; i++) {
        const int64_t xlen = offsets[i + 1] - offsets[i];
        const or_hit *x = hits + offsets[i];
        if (ray_len_out) ray_len_out[i] = rays[i].length;
        if (ray_tm_out) ray_tm_out[i] = rays[i].Nph;
        if (xlen == 0 || rays[i].Nph == 0) continue;
        double TM = rays[i].Nph;
        if (f->no_photon && rays[i].type < 0) TM = 0.0;
        const double freq = rays[i].freq;
        const double sigmaHI = or_xs_get(c, 0, freq);
        const double sigmaHeI = or_xs_get(c, 1, freq);
        const double sigmaHeII = or_xs_get(c, 2, freq);
        const double dEHI = fdim_(freq, u->C_HI_ENERGY);
        const double dEHeI = fdim_(freq, u->C_HEI_ENERGY);
        const double dEHeII = fdim_(freq, u->C_HEII_ENERGY);
        const double hf_HI = u->C_H_PER_MASS * dEHI;
        const double hf_HeI = u->C_HE_PER_MASS * dEHeI;
        const double hf_HeII = u->C_HE_PER_MASS * dEHeII;
        total += xlen;
        int64_t j;
        for (j = 0; j < xlen; j++) {
            const int64_t ip = x[j].ipar;
            const double b = x[j].b;
            const double sml = p->sml[ip];
            const double sml_inv = 1.0 / sml;
            double NH = p->mass[ip] * u->C_H_PER_MASS;
            double NHe = p->mass[ip] * u->C_HE_PER_MASS;
            const double NH_inv = 1.0 / NH;
            const double NHe_inv = 1.0 / NHe;
            const double xHI = p->lambdaH[ip];
            const double xHeI = p->lambdaHeI[ip];
            const double xHeII = p->lambdaHeII[ip];
            const double kernel = or_sph_depth(c, b * sml_inv) * (sml_inv * sml_inv) * scaling_fac2_inv;
            double deltaHI = 0.0, deltaHeI = 0.0, deltaHeII = 0.0;
            double tausum = 0.0, tauHI = 0.0, tauHeI = 0.0, tauHeII = 0.0;
            const double NHI = xHI * NH;
            const double NHIcd = kernel * NHI;
            tauHI = sigmaHI * NHIcd;
            tausum += tauHI;
            const int he = !f->h_only && NHe != 0.0;
            if (he) {
                const double NHeIcd = kernel * (xHeI * NHe);
                tauHeI = sigmaHeI * NHeIcd;
                tausum += tauHeI;
                const double NHeIIcd = kernel * (xHeII * NHe);
                tauHeII = sigmaHeII * NHeIIcd;
                tausum += tauHeII;
            }
            if (tausum == 0.0) continue;
            const double absorb_est = -TM * expm1(-tausum);
            double absorb = 0;
            double absorbHI = tauHI / tausum * absorb_est;
            if (absorbHI > NHI) {
                absorbHI = NHI;
                saturated++;
            }
            deltaHI = absorbHI * NH_inv;
            absorb = absorbHI;
            if (he) {
                const double NHeI = xHeI * NHe;
                const double NHeII = xHeII * NHe;
                double absorbHeI = tauHeI / tausum * absorb_est;
                if (absorbHeI > NHeI) {
                    absorbHeI = NHeI;
                    saturated++;
                }
                deltaHeI = absorbHeI * NHe_inv;
                double absorbHeII = tauHeII / tausum * absorb_est;
                if (absorbHeII > NHeII) {
                    absorbHeII = NHeII;
                    saturated++;
                }
                deltaHeII = absorbHeII * NHe_inv;
                absorb += (absorbHeI + absorbHeII);
            }
            double xe = psys_ye(c, p, ip) * NH / (NH + NHe);
            if (xe > 1.0) xe = 1.0;
            if (f->secondary_ionization && xe > 0.0) {
                add_f32(&p->yGdepHI[ip], deltaHI, atomic);
                fiHI += deltaHI * NH;
                const double log10x = log10(xe);
                /* the HeII terms after the stray ';' are dropped, src/loop.c:530-531 */
                const double secion =
                    deltaHI * dEHI / u->C_HI_ENERGY * or_secion_get(c, 0, dEHI, log10x) +
                    deltaHeI * dEHeI / u->C_HI_ENERGY * or_secion_get(c, 0, dEHeI, log10x);
                add_f32(&p->yGdepHI[ip], secion, atomic);
                siHI += secion * NH;
                add_f32(&p->heat[ip], hf_HI * deltaHI * or_secion_get(c, 2, dEHI, log10x), atomic);
                if (he) {
                    const double secion2 =
                        deltaHI * dEHI / u->C_HEI_ENERGY * or_secion_get(c, 1, dEHI, log10x) +
                        deltaHeI * dEHeI / u->C_HEI_ENERGY * or_secion_get(c, 1, dEHeI, log10x);
                    add_f32(&p->yGdepHeI[ip], deltaHeI + secion2, atomic);
                    siHeI += secion2 * NHe;
                    fiHeI += deltaHeI * NHe;
                    add_f32(&p->heat[ip], hf_HeI * deltaHeI * or_secion_get(c, 2, dEHeI, log10x), atomic);
                    add_f32(&p->yGdepHeII[ip], deltaHeII, atomic);
                    fiHeII += deltaHeII * NHe;
                }
            } else {
                add_f32(&p->yGdepHI[ip], deltaHI, atomic);
                add_f32(&p->heat[ip], hf_HI * deltaHI, atomic);
                if (he) {
                    add_f32(&p->yGdepHeI[ip], deltaHeI, atomic);
                    add_f32(&p->yGdepHeII[ip], deltaHeII, atomic);
                    add_f32(&p->heat[ip], deltaHeI * hf_HeI, atomic);
                    add_f32(&p->heat[ip], deltaHeII * hf_HeII, atomic);
                }
            }
            add_i32(&p->hits[ip], atomic);
            TM -= absorb;
            if (TM < 0.0 || TM / rays[i].Nph < 1e-10) {
                j++;
                break;
            }
        }
        if (ray_len_out) ray_len_out[i] = x[j - 1].d;
        if (ray_tm_out) ray_tm_out[i] = TM;
        lost += TM;
    }
    if (st) {
        st->deposit_time += now_s() - t0;
        st->total_deposit_count += total;
        st->saturated_deposit_count += saturated;
        st->lost_photon_count_sum += lost;
        st->first_ionization_HI += fiHI;
        st->first_ionization_HeI += fiHeI;
        st->first_ionization_HeII += fiHeII;
        st->secondary_ionization_HI += siHI;
        st->secondary_ionization_HeI += siHeI;
    }
    return 0;
}

/* ----------------------------------------------------------------- merge */
/* src/loop.c:610-645 */
int64_t or_merge(const or_psys *p, int64_t nray, const int64_t *offsets, const or_hit *hits,
                 int64_t *active) {
    uint8_t *mask = calloc(p->npar > 0 ? p->npar : 1, 1);
    for (int64_t k = 0; k < offsets[nray]; k++) mask[hits[k].ipar] = 1;
    int64_t n = 0;
    for (int64_t k = 0; k < offsets[nray]; k++) {
        const int64_t ip = hits[k].ipar;
        if (mask[ip]) {
            active[n++] = ip;
            mask[ip] = 0;
        }
    }
    free(mask);
    return n;
}

/* -------------------------------------------------------------- chemistry */
double or_ieye2T(const or_ctx *c, double ie, double ye) {
    const or_units *u = &c->u;
    return u->U_MPROTON_OVER_C_BOLTZMANN * ie / (u->C_HEMF * 0.25 + u->C_HMF + u->C_HMF * ye) *
           2.0 / 3.0;
}
double or_Tye2ie(const or_ctx *c, double T, double ye) {
    const or_units *u = &c->u;
    return T * u->C_BOLTZMANN_OVER_U_MPROTON * (u->C_HEMF * 0.25 + u->C_HMF + u->C_HMF * ye) /
           2.0 * 3.0;
}

/* src/solve.c:42-141 (FETCH_VARS + function) */
int or_chem_rhs(double t, double *x, double *dxdt, void *data) {
    (void)t;
    const or_chem_data *cd = data;
    const or_ctx *c = cd->c;
    const or_step *step = cd->step;
    const or_flags *f = &c->f;
    const or_units *u = &c->u;
    const double xHI = x[0];
    const double xHII = 1.0 - x[0];
    const double xHeI = x[1];
    const double xHeII = x[2];
    const double xHeIII = 1.0 - x[2] - x[1];
    const double ye = (1.0 - x[0]) + u->C_HEMF / u->C_HMF * 0.25 * (x[2] + 2 * (1.0 - x[2] - x[1])) +
                      step->yeMET;
    const double nH = step->nH;
    const double T = f->isothermal ? step->T : or_ieye2T(c, x[6], ye);
    const double logT = log10(T);
    const double gamma_HI = or_ar_get(c, AR_HI_CI, logT) * nH;
    const double alpha_HII_A = or_ar_get(c, AR_HII_RC_A, logT) * nH;
    const double alpha_HII_B = or_ar_get(c, AR_HII_RC_B, logT) * nH;
    const double alpha_HII_AB = fdim_(alpha_HII_A, alpha_HII_B);
    const double gamma_HeI = or_ar_get(c, AR_HEI_CI, logT) * nH;
    const double gamma_HeII = or_ar_get(c, AR_HEII_CI, logT) * nH;
    const double alpha_HeII_A = or_ar_get(c, AR_HEII_RC_A, logT) * nH;
    const double alpha_HeII_B = or_ar_get(c, AR_HEII_RC_B, logT) * nH;
    const double alpha_HeII_AB = fdim_(alpha_HeII_A, alpha_HeII_B);
    const double alpha_HeIII_A = or_ar_get(c, AR_HEIII_RC_A, logT) * nH;
    const double alpha_HeIII_B = or_ar_get(c, AR_HEIII_RC_B, logT) * nH;
    const double alpha_HeIII_AB = fdim_(alpha_HeIII_A, alpha_HeIII_B);
    const double eta_HII =
        (f->on_the_spot ? or_ar_get(c, AR_HII_RCC_B, logT) : or_ar_get(c, AR_HII_RCC_A, logT)) * nH;
    const double psi_HI = or_ar_get(c, AR_HI_CEC, logT) * nH;
    const double zeta_HI = or_ar_get(c, AR_HI_CIC, logT) * nH;
    const double beta = or_ar_get(c, AR_E_BREMC, logT) * nH;
    const double chi = or_ar_get(c, AR_E_COMPC, logT);
    if (!f->on_the_spot) {
        dxdt[0] = -step->yGdepHI / step->time + (-gamma_HI * ye * xHI + alpha_HII_A * ye * xHII);
        dxdt[3] = alpha_HII_AB * ye * xHII;
        if (f->h_only) {
            dxdt[1] = dxdt[2] = dxdt[4] = dxdt[5] = 0.0;
        } else {
            dxdt[1] = -step->yGdepHeI / step->time + (-gamma_HeI * ye * xHeI + alpha_HeII_A * ye * xHeII);
            dxdt[2] = -step->yGdepHeII / step->time +
                      (-gamma_HeII * ye * xHeII - alpha_HeII_A * ye * xHeII + alpha_HeIII_A * ye * xHeIII);
            dxdt[4] = alpha_HeII_AB * ye * xHeII;
            dxdt[5] = alpha_HeIII_AB * ye * xHeIII;
        }
    } else {
        dxdt[0] = -step->yGdepHI / step->time + (-gamma_HI * ye * xHI + alpha_HII_B * ye * xHII);
        if (f->h_only) {
            dxdt[1] = dxdt[2] = 0.0;
        } else {
            /* xHI (not xHeI) as in src/solve.c:103-104 */
            dxdt[1] = -step->yGdepHeI / step->time + (-gamma_HeI * ye * xHI + alpha_HeII_B * ye * xHeII);
            dxdt[2] = -step->yGdepHeII / step->time +
                      (-gamma_HeII * ye * xHI - alpha_HeII_B * ye * xHeII + alpha_HeIII_B * ye * xHeIII);
        }
        dxdt[3] = dxdt[4] = dxdt[5] = 0;
    }
    if (f->adiabatic | f->isothermal) {
        dxdt[6] = 0.0;
    } else {
        const double L = u->C_H_PER_MASS * ((zeta_HI + psi_HI) * ye * xHI + (eta_HII * ye * xHII) +
                                            beta * ye * xHII + chi * ye);
        dxdt[6] = step->heat / step->time - L;
    }
    for (int i = 0; i < 7; i++) dxdt[i] *= step->time;
    return 0;
}

/* ---- the right-hand side with its Jacobian by forward-mode dual numbers:
 * values with the operations and order of or_chem_rhs, derivatives with
 * respect to the 7 variables; the rates' derivative in log10 T is the
 * interpolation segment's slope (the same algorithm as the device). */
typedef struct { double v, d[7]; } dual7;
static dual7 dc(double v) { dual7 r; r.v = v; for (int k = 0; k < 7; k++) r.d[k] = 0.0; return r; }
static dual7 dadd(dual7 a, dual7 b) { dual7 r; r.v = a.v + b.v; for (int k = 0; k < 7; k++) r.d[k] = a.d[k] + b.d[k]; return r; }
static dual7 dsub(dual7 a, dual7 b) { dual7 r; r.v = a.v - b.v; for (int k = 0; k < 7; k++) r.d[k] = a.d[k] - b.d[k]; return r; }
static dual7 dmul(dual7 a, dual7 b) {
    dual7 r;
    r.v = a.v * b.v;
    for (int k = 0; k < 7; k++) r.d[k] = a.d[k] * b.v + a.v * b.d[k];
    return r;
}
static dual7 dneg(dual7 a) { dual7 r; r.v = -a.v; for (int k = 0; k < 7; k++) r.d[k] = -a.d[k]; return r; }
static dual7 cadd(double a, dual7 b) { dual7 r = b; r.v = a + b.v; return r; }    /* a + b */
static dual7 addc(dual7 a, double b) { dual7 r = a; r.v = a.v + b; return r; }    /* a + b */
static dual7 csub(double a, dual7 b) { dual7 r; r.v = a - b.v; for (int k = 0; k < 7; k++) r.d[k] = -b.d[k]; return r; }
static dual7 cmul(double a, dual7 b) { dual7 r; r.v = a * b.v; for (int k = 0; k < 7; k++) r.d[k] = a * b.d[k]; return r; }
static dual7 dfdim(dual7 a, dual7 b) { return a.v > b.v ? dsub(a, b) : dc(fdim_(a.v, b.v)); }

/* tab_get and its slope in value */
static double tab_get_slope(const or_tab *t, int id, double value, double *slope) {
    ssize_t index = (ssize_t)((value - t->min) * t->step_inv);
    *slope = 0.0;
    if (index < 0) return t->data[id][0];
    if (index >= t->nrows - 1) return t->data[id][t->nrows - 1];
    double left = t->data[0][index];
    double right = t->data[0][index + 1];
    double leftwt = (right - value);
    double rightwt = (value - left);
    *slope = (t->data[id][index + 1] - t->data[id][index]) * t->step_inv;
    return (leftwt * t->data[id][index] + rightwt * t->data[id][index + 1]) * t->step_inv;
}

/* rate col at log10 T times nh, with its derivatives through dual L */
static dual7 drate(const or_ctx *c, int col, dual7 L, double nh, int iso) {
    double sl;
    dual7 r;
    r.v = tab_get_slope(&c->ar, c->arcol[col], L.v, &sl) * nh;
    for (int k = 0; k < 7; k++) r.d[k] = iso ? 0.0 : sl * nh * L.d[k];
    return r;
}

/* f = or_chem_rhs(x) and jac[i][j] = d f_i / d x_j (7 x 7, row-major) */
int or_chem_jac(const or_ctx *c, const or_step *step, const double *x, double *f, double *jac) {
    const or_flags *fl = &c->f;
    const or_units *u = &c->u;
    dual7 X[7];
    for (int q = 0; q < 7; q++) { X[q] = dc(x[q]); X[q].d[q] = 1.0; }
    const dual7 xHI = X[0];
    const dual7 xHII = csub(1.0, X[0]);
    const dual7 xHeI = X[1];
    const dual7 xHeII = X[2];
    const dual7 xHeIII = dsub(csub(1.0, X[2]), X[1]);
    const dual7 ye = addc(dadd(csub(1.0, X[0]), cmul(u->C_HEMF / u->C_HMF * 0.25,
                                                     dadd(X[2], cmul(2, dsub(csub(1.0, X[2]), X[1]))))),
                          step->yeMET);
    const double nH = step->nH;
    const int iso = fl->isothermal != 0;
    dual7 L = dc(0.0);
    if (iso) {
        L.v = log10(step->T);
    } else {
        const double den = u->C_HEMF * 0.25 + u->C_HMF + u->C_HMF * ye.v;
        L.v = log10(or_ieye2T(c, x[6], ye.v));
        const double dye = -u->C_HMF * 0.43429448190325182765 / den;
        for (int k = 0; k < 7; k++) L.d[k] = dye * ye.d[k];
        L.d[6] += 0.43429448190325182765 / x[6];
    }
    const dual7 gamma_HI = drate(c, AR_HI_CI, L, nH, iso);
    const dual7 gamma_HeI = drate(c, AR_HEI_CI, L, nH, iso);
    const dual7 gamma_HeII = drate(c, AR_HEII_CI, L, nH, iso);
    dual7 g[7];
    /* -a * ye * b as ((-a) * ye) * b */
#define NAB(a, b) dmul(dmul(dneg(a), ye), (b))
#define AB(a, b) dmul(dmul((a), ye), (b))
    if (!fl->on_the_spot) {
        const dual7 aHA = drate(c, AR_HII_RC_A, L, nH, iso), aHB = drate(c, AR_HII_RC_B, L, nH, iso);
        g[0] = cadd(-step->yGdepHI / step->time, dadd(NAB(gamma_HI, xHI), AB(aHA, xHII)));
        g[3] = AB(dfdim(aHA, aHB), xHII);
        if (fl->h_only) {
            g[1] = g[2] = g[4] = g[5] = dc(0.0);
        } else {
            const dual7 a2A = drate(c, AR_HEII_RC_A, L, nH, iso), a2B = drate(c, AR_HEII_RC_B, L, nH, iso);
            const dual7 a3A = drate(c, AR_HEIII_RC_A, L, nH, iso), a3B = drate(c, AR_HEIII_RC_B, L, nH, iso);
            g[1] = cadd(-step->yGdepHeI / step->time, dadd(NAB(gamma_HeI, xHeI), AB(a2A, xHeII)));
            g[2] = cadd(-step->yGdepHeII / step->time,
                        dadd(dsub(NAB(gamma_HeII, xHeII), AB(a2A, xHeII)), AB(a3A, xHeIII)));
            g[4] = AB(dfdim(a2A, a2B), xHeII);
            g[5] = AB(dfdim(a3A, a3B), xHeIII);
        }
    } else {
        const dual7 aHB = drate(c, AR_HII_RC_B, L, nH, iso);
        g[0] = cadd(-step->yGdepHI / step->time, dadd(NAB(gamma_HI, xHI), AB(aHB, xHII)));
        if (fl->h_only) {
            g[1] = g[2] = dc(0.0);
        } else {
            const dual7 a2B = drate(c, AR_HEII_RC_B, L, nH, iso), a3B = drate(c, AR_HEIII_RC_B, L, nH, iso);
            g[1] = cadd(-step->yGdepHeI / step->time, dadd(NAB(gamma_HeI, xHI), AB(a2B, xHeII)));
            g[2] = cadd(-step->yGdepHeII / step->time,
                        dadd(dsub(NAB(gamma_HeII, xHI), AB(a2B, xHeII)), AB(a3B, xHeIII)));
        }
        g[3] = g[4] = g[5] = dc(0.0);
    }
    if (fl->adiabatic | fl->isothermal) {
        g[6] = dc(0.0);
    } else {
        const dual7 zeta = drate(c, AR_HI_CIC, L, nH, iso), psi = drate(c, AR_HI_CEC, L, nH, iso);
        const dual7 eta = drate(c, fl->on_the_spot ? AR_HII_RCC_B : AR_HII_RCC_A, L, nH, iso);
        const dual7 beta = drate(c, AR_E_BREMC, L, nH, iso);
        const dual7 chi = drate(c, AR_E_COMPC, L, 1.0, iso);
        const dual7 Lc = cmul(u->C_H_PER_MASS,
                              dadd(dadd(dadd(AB(dadd(zeta, psi), xHI), AB(eta, xHII)), AB(beta, xHII)),
                                   dmul(chi, ye)));
        g[6] = csub(step->heat / step->time, Lc);
    }
#undef NAB
#undef AB
    for (int i = 0; i < 7; i++) {
        f[i] = g[i].v * step->time;
        if (jac)
            for (int j = 0; j < 7; j++) jac[i * 7 + j] = g[i].d[j] * step->time;
    }
    return 0;
}

/* variables the flags leave evolving (the others have zero derivative) */
static int chem_active(const or_flags *f, int act[7]) {
    int n = 0;
    act[n++] = 0;
    if (!f->h_only) { act[n++] = 1; act[n++] = 2; }
    if (!f->on_the_spot) {
        act[n++] = 3;
        if (!f->h_only) { act[n++] = 4; act[n++] = 5; }
    }
    if (!(f->adiabatic || f->isothermal)) act[n++] = 6;
    return n;
}

/* LU with partial pivoting on an n x n row-major matrix (n <= 7) */
static int lu_factor(double *a, int n, int *piv) {
    for (int k = 0; k < n; k++) {
        int p = k;
        double m = fabs(a[k * n + k]);
        for (int i = k + 1; i < n; i++)
            if (fabs(a[i * n + k]) > m) { m = fabs(a[i * n + k]); p = i; }
        piv[k] = p;
        if (m == 0.0) return -1;
        if (p != k)
            for (int j = 0; j < n; j++) {
                double tmp = a[k * n + j];
                a[k * n + j] = a[p * n + j];
                a[p * n + j] = tmp;
            }
        const double inv = 1.0 / a[k * n + k];
        a[k * n + k] = inv; /* the solves multiply by the pivot's reciprocal */
        for (int i = k + 1; i < n; i++) {
            const double l = a[i * n + k] * inv;
            a[i * n + k] = l;
            for (int j = k + 1; j < n; j++) a[i * n + j] -= l * a[k * n + j];
        }
    }
    return 0;
}
static void lu_solve(const double *a, int n, const int *piv, double *b) {
    for (int k = 0; k < n; k++) {
        if (piv[k] != k) {
            double tmp = b[k];
            b[k] = b[piv[k]];
            b[piv[k]] = tmp;
        }
        for (int i = k + 1; i < n; i++) b[i] -= a[i * n + k] * b[k];
    }
    for (int k = n - 1; k >= 0; k--) {
        double s = b[k];
        for (int j = k + 1; j < n; j++) s -= a[k * n + j] * b[j];
        b[k] = s * a[k * n + k];
    }
}

/* ROS3 of Sandu, Verwer, Blom, Spee, Carmichael & Potra, "Benchmarking stiff
 * ODE solvers for atmospheric chemistry problems II: Rosenbrock solvers",
 * Atmos. Environ. 31 (1997) 3459-3472: three stages, two evaluations of F,
 * order 3 with an embedded order-2 solution, L-stable.  Coefficients of the
 * form G = I/(gamma h) - J, G K_i = F(y + sum a_ij K_j) + sum (c_ij/h) K_j. */
#define ROS3_GAMMA 0.43586652150845899941601945119356
#define ROS3_C21 (-1.0156171083877702091975600115545)
#define ROS3_C31 4.0759956452537699824805835358067
#define ROS3_C32 9.2076794298330791242156818474003
#define ROS3_M1 1.0
#define ROS3_M2 6.1697947043828245592553615689730
#define ROS3_M3 (-0.42772256543218573326238373806514)
#define ROS3_E1 0.5
#define ROS3_E2 (-2.9079558716805469821718236208017)
#define ROS3_E3 0.22354069897811569627360909276199
#define ROS3_MAXSTEPS 100000

/* v^(-1/3) for the step-size controller, v in [1e-10, 1e10]: with v = m 2^e,
 * m in [0.5, 1), e = 3q + r, v^(-1/3) = m^(-1/3) 2^(-r/3) 2^(-q), and
 * m^(-1/3) by Newton's iteration for the reciprocal cube root, y += y (1 -
 * m y^3) / 3 (no division; 4 steps from the linear guess reach 1 ulp).  Only
 * correctly rounded operations, so the device computes the same bits */
static double ctl_rcbrt(double v) {
    int e;
    const double m = frexp(v, &e);
    const int r = ((e % 3) + 3) % 3;
    const int q = (e - r) / 3;
    double y = 1.4 - 0.4 * m;
    for (int k = 0; k < 4; k++) y = y + y * (1.0 - m * (y * y * y)) * (1.0 / 3.0);
    const double c = r == 0 ? 1.0 : (r == 1 ? 0.79370052598409973738 : 0.62996052494743658238);
    return ldexp(y * c, -q);
}

/* Isothermal, hydrogen-only, on-the-spot: the rates are fixed over the step
 * and src/solve.c:97-99 reduces to the Riccati equation (x = x_HI, tau = t /
 * time, G = gamma_HI time, A = alpha_B time, g = yGdepHI, m = yeMET)
 *     dx/dtau = F(x) = -g + (1 + m - x)(A - c x),   c = A + G,
 * i.e. F = a + b x + c x^2 with a = (1+m)A - g, b = -((1+m)c + A).  Its roots
 * are real, lambda^2 = b^2 - 4ac = ((1+m)G + mA)^2 + 4cg >= 0, r1 = a/q <= r2
 * = q/c with q = (lambda - b)/2.  With u = x - r1 the solution is
 *     u(1) = u0 e^-lambda / (e^-lambda + w (1 - e^-lambda)),  w = c (r2 - x0) / lambda,
 * evaluated as x(1) = x0 + F(x0) phi / D, phi = (1 - e^-lambda) / lambda,
 * D = e^-lambda + w (1 - e^-lambda); w is taken from F(x0) = -c u0 (r2 - x0)
 * when x0 is nearer r2 (no cancellation in r2 - x0).  Exact in one step: it
 * replaces the integrator for this flag set (the state is still clamped). */
double or_riccati_tick(double x0, double g, double m, double G, double A) {
    const double c = A + G;
    const double F0 = -g + (1.0 + m - x0) * (A - c * x0);
    if (F0 == 0.0) return x0;
    const double b = -((1.0 + m) * c + A);
    const double a = (1.0 + m) * A - g;
    const double dd = (1.0 + m) * G + m * A;
    const double lam = sqrt(dd * dd + 4.0 * c * g);
    const double q = 0.5 * (lam - b);
    if (lam < 1e-8) {
        /* F nearly constant: phi = 1 - lambda/2 + O(lambda^2), D = 1 - c u0 phi */
        const double cu0 = q > 0 ? c * x0 - a * c / q : 0.0;
        const double phi = 1.0 - 0.5 * lam;
        return x0 + F0 * phi / (1.0 - cu0 * phi);
    }
    const double u0 = x0 - a / q;
    const double s = -expm1(-lam);
    const double w = (c * u0 > 0.5 * lam) ? -F0 / (u0 * lam) : (q - c * x0) / lam;
    const double D = (1.0 - s) + fmax(w, 0.0) * s;
    return x0 + F0 * (s / lam) / D;
}

/* One ROS3 attempt of size h from x with F(x) = f0 and the Jacobian J over
 * the n evolving variables act[]: the new state xn[7] and the embedded
 * error err[n]; returns -1 when I/(gamma h) - J is singular. */
static int ros3_attempt(or_chem_data *cd, const int *act, int n, const double *x, const double *f0,
                        const double *J, double t, double h, double *xn, double *err) {
    double M[49], f1[7], xt[7], k1[7], k2[7], k3[7];
    int piv[7];
    const double ghinv = 1.0 / (ROS3_GAMMA * h);
    for (int ii = 0; ii < n; ii++)
        for (int jj = 0; jj < n; jj++) M[ii * n + jj] = (ii == jj ? ghinv : 0.0) - J[ii * n + jj];
    if (lu_factor(M, n, piv)) return -1;
    const double c21 = ROS3_C21 / h, c31 = ROS3_C31 / h, c32 = ROS3_C32 / h;
    for (int ii = 0; ii < n; ii++) k1[ii] = f0[act[ii]];
    lu_solve(M, n, piv, k1);
    memcpy(xt, x, sizeof(xt));
    for (int ii = 0; ii < n; ii++) xt[act[ii]] = x[act[ii]] + k1[ii];
    or_chem_rhs(t + h, xt, f1, cd);
    for (int ii = 0; ii < n; ii++) k2[ii] = f1[act[ii]] + c21 * k1[ii];
    lu_solve(M, n, piv, k2);
    for (int ii = 0; ii < n; ii++) k3[ii] = f1[act[ii]] + c31 * k1[ii] + c32 * k2[ii];
    lu_solve(M, n, piv, k3);
    memcpy(xn, x, sizeof(double) * 7);
    for (int ii = 0; ii < n; ii++) {
        const int i = act[ii];
        xn[i] = x[i] + ROS3_M1 * k1[ii] + ROS3_M2 * k2[ii] + ROS3_M3 * k3[ii];
        err[ii] = ROS3_E1 * k1[ii] + ROS3_E2 * k2[ii] + ROS3_E3 * k3[ii];
    }
    return 0;
}

static void jac_active(const or_ctx *c, const or_step *s, const int *act, int n, const double *x,
                       double *f0, double *J) {
    double J7[49];
    or_chem_jac(c, s, x, f0, J7);
    for (int ii = 0; ii < n; ii++)
        for (int jj = 0; jj < n; jj++) J[ii * n + jj] = J7[act[ii] * 7 + act[jj]];
}

/* Three-stage L-stable Rosenbrock method ROS3 (coefficients above):
 *   G = I/(gamma h) - J,   G K1 = F(y),   G K2 = F(y + K1) + (c21/h) K1,
 *   G K3 = F(y + K1) + (c31/h) K1 + (c32/h) K2,
 *   y' = y + m1 K1 + m2 K2 + m3 K3,   err = e1 K1 + e2 K2 + e3 K3,
 * step factor 0.9 E^(-1/3) within [0.2, 5], E the RMS of err over the
 * evolving variables scaled by atol + rtol max(|y|, |y'|).  Exact Jacobian
 * over the evolving variables (or_chem_jac, dual numbers).  Tolerances of
 * src/solve.c:151-152 unless overridden. */
int32_t or_step_evolve(const or_ctx *c, or_step *s, double rtol, const double *atol7) {
    static const double atol_def[7] = {1e-7, 1e-7, 1e-7, 1e-7, 1e-7, 1e-7, 1.};
    const double *atol = atol7 ? atol7 : atol_def;
    if (rtol <= 0) rtol = 1e-4;
    or_chem_data cd = {c, s};
    int act[7];
    const int n = chem_active(&c->f, act);
    double x[7] = {s->lambdaH, s->lambdaHeI, s->lambdaHeII, s->yGrecHII, s->yGrecHeII,
                   s->yGrecHeIII, s->ie};
    double t = 0.0, h = 1.0;
    int nsteps = 0, rejected = 0;
    double f0[7], J[49], err[7];
    if (c->f.isothermal && c->f.h_only && c->f.on_the_spot) {
        const double logT = log10(s->T);
        const double G = or_ar_get(c, AR_HI_CI, logT) * s->nH * s->time;
        const double A = or_ar_get(c, AR_HII_RC_B, logT) * s->nH * s->time;
        x[0] = or_riccati_tick(x[0], s->yGdepHI, s->yeMET, G, A);
        nsteps = 1;
        t = 1.0;
    }
    int dead = 0;
    while (!dead && t < 1.0 && nsteps < ROS3_MAXSTEPS) {
        if (h > 1.0 - t) h = 1.0 - t;
        jac_active(c, s, act, n, x, f0, J);
        for (;;) {
            nsteps++;
            double E = INFINITY;
            double xn[7];
            if (!ros3_attempt(&cd, act, n, x, f0, J, t, h, xn, err)) {
                double acc = 0.0;
                for (int ii = 0; ii < n; ii++) {
                    const int i = act[ii];
                    const double sc = atol[i] + rtol * fmax(fabs(x[i]), fabs(xn[i]));
                    acc += (err[ii] / sc) * (err[ii] / sc);
                }
                E = sqrt(acc / n);
                if (!(E == E)) E = INFINITY;
            }
            /* a singular matrix or a non-finite error estimate never
             * advances the state; at the smallest step it ends the
             * integration with x unchanged (a solver failure, t < 1) */
            const int good = E < INFINITY;
            if (!good && h < 1e-14) {
                dead = 1;
                break;
            }
            if (good && (E <= 1.0 || h < 1e-14)) {
                t += h;
                memcpy(x, xn, sizeof(x));
                double fac = 0.9 * ctl_rcbrt(fmin(fmax(E, 1e-10), 1e10));
                fac = fmin(5.0, fmax(0.2, fac));
                if (rejected) fac = fmin(fac, 1.0);
                rejected = 0;
                h *= fac;
                if (1.0 - t < 1e-13) t = 1.0;
                break;
            }
            rejected = 1;
            double fac = (E == INFINITY) ? 0.2 : fmax(0.2, 0.9 * ctl_rcbrt(fmin(E, 1e10)));
            h *= fac;
        }
    }
    /* clamps and write-back, src/solve.c:208-223 */
    x[0] = fmin(1., fmax(0., x[0]));
    x[1] = fmin(1., fmax(0., x[1]));
    x[2] = fmin(1. - x[1], fmax(0., x[2]));
    x[3] = fmax(0., x[3]);
    x[4] = fmax(0., x[4]);
    x[5] = fmax(0., x[5]);
    s->lambdaH = x[0];
    s->lambdaHeI = x[1];
    s->lambdaHeII = x[2];
    s->yGrecHII = x[3];
    s->yGrecHeII = x[4];
    s->yGrecHeIII = x[5];
    s->ie = x[6];
    s->step_remain = 0.0;
    s->nsteps = nsteps;
    return t >= 1.0 ? 0 : -1;
}

/* ROS3 with nsub equal steps and no step control, for the convergence-order
 * test: the final state x7 (unclamped) and the sum of the embedded error
 * vectors err[7] (zero in frozen variables); -1 on a singular matrix */
int32_t or_ros3_fixed(const or_ctx *c, const or_step *s, int32_t nsub, double *x7, double *err7) {
    or_chem_data cd = {c, s};
    int act[7];
    const int n = chem_active(&c->f, act);
    double x[7] = {s->lambdaH, s->lambdaHeI, s->lambdaHeII, s->yGrecHII, s->yGrecHeII,
                   s->yGrecHeIII, s->ie};
    double f0[7], J[49], err[7], xn[7];
    const double h = 1.0 / nsub;
    for (int q = 0; q < 7; q++) err7[q] = 0.0;
    for (int k = 0; k < nsub; k++) {
        jac_active(c, s, act, n, x, f0, J);
        if (ros3_attempt(&cd, act, n, x, f0, J, k * h, h, xn, err)) return -1;
        memcpy(x, xn, sizeof(x));
        for (int ii = 0; ii < n; ii++) err7[act[ii]] += err[ii];
    }
    memcpy(x7, x, sizeof(x));
    return 0;
}

int32_t or_step_evolve_many(const or_ctx *c, or_step *s, int64_t n, double rtol,
                            int32_t nthreads) {
    int32_t bad = 0;
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = 1;
#pragma omp parallel for schedule(dynamic, 64) num_threads(nthreads) reduction(+ : bad)
#endif
    for (int64_t i = 0; i < n; i++) bad += or_step_evolve(c, &s[i], rtol, NULL) != 0;
    return bad;
}

/* src/loop.c:647-776 (EOS disabled; the LSODA path leaves step_remain 0) */
int32_t or_update(or_ctx *c, or_psys *p, const int64_t *active, int64_t nactive, int64_t tick,
                  double tick_time, or_stats *st, int32_t nthreads) {
    const or_flags *f = &c->f;
    const or_units *u = &c->u;
    const double scaling_fac = f->comoving ? 1 / (f->redshift + 1.0) : 1.0;
    const double scaling_fac3_inv = 1.0 / (scaling_fac * scaling_fac * scaling_fac);
    int64_t steps = 0, bad = 0;
    double t0 = now_s();
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = 1;
#pragma omp parallel for schedule(dynamic, 64) num_threads(nthreads) reduction(+ : steps, bad)
#endif
    for (int64_t k = 0; k < nactive; k++) {
        const int64_t ip = active[k];
        const double rho = p->rho[ip] * scaling_fac3_inv;
        or_step s;
        memset(&s, 0, sizeof(s));
        s.ipar = ip;
        s.time = (tick - p->lasthit[ip]) * tick_time;
        if (f->fake_temperature > 0)
            s.T = f->fake_temperature;
        else
            s.T = or_ieye2T(c, p->ie[ip], psys_ye(c, p, ip));
        s.yGdepHI = p->yGdepHI[ip];
        s.yGdepHeI = p->yGdepHeI[ip];
        s.yGdepHeII = p->yGdepHeII[ip];
        s.heat = p->heat[ip];
        s.yGrecHII = p->yGrecHII[ip];
        s.yGrecHeII = p->yGrecHeII[ip];
        s.yGrecHeIII = p->yGrecHeIII[ip];
        s.lambdaH = p->lambdaH[ip];
        s.lambdaHeI = p->lambdaHeI[ip];
        s.lambdaHeII = p->lambdaHeII[ip];
        s.yeMET = p->yeMET[ip];
        s.nH = u->C_H_PER_MASS * rho;
        s.ie = p->ie[ip];
        bad += or_step_evolve(c, &s, 1e-4, NULL) != 0;
        steps += s.nsteps;
        p->yGdepHI[ip] = (float)(p->yGdepHI[ip] * s.step_remain);
        p->yGdepHeI[ip] = (float)(p->yGdepHeI[ip] * s.step_remain);
        p->yGdepHeII[ip] = (float)(p->yGdepHeII[ip] * s.step_remain);
        p->yGrecHII[ip] = (float)s.yGrecHII;
        p->yGrecHeII[ip] = (float)s.yGrecHeII;
        p->yGrecHeIII[ip] = (float)s.yGrecHeIII;
        p->lambdaH[ip] = fmax(0, fmin(1, s.lambdaH));
        p->lambdaHeI[ip] = fmax(0, fmin(1, s.lambdaHeI));
        p->lambdaHeII[ip] = fmax(0, fmin(1 - p->lambdaHeI[ip], s.lambdaHeII));
        if (!f->adiabatic) p->ie[ip] = (float)s.ie;
        p->lasthit[ip] = tick;
        if (f->isothermal) p->ie[ip] = (float)or_Tye2ie(c, s.T, psys_ye(c, p, ip));
        p->heat[ip] = 0.0f;
    }
    if (st) {
        st->update_time += now_s() - t0;
        st->evolve_count += nactive;
        st->solver_steps += steps;
    }
    return bad ? -1 : 0;
}

/* src/loop.c:131-148 (emit_rays is replaced by the given ray list) */
int32_t or_tick(or_ctx *c, const or_tree *t, or_psys *p, const or_ray *rays, int64_t nray,
                int64_t tick, double tick_time, or_stats *st, int32_t nthreads) {
    int64_t *offsets = malloc(sizeof(int64_t) * (nray + 1));
    double t0 = now_s();
    or_hit *hits = trace_all(t, p, rays, nray, offsets, nthreads);
    const int64_t total = offsets[nray];
    if (st) st->trace_time += now_s() - t0;
    if (!c->f.trace_only) or_deposit(c, p, rays, nray, offsets, hits, NULL, NULL, st, nthreads);
    t0 = now_s();
    int64_t *active = malloc(sizeof(int64_t) * (total > 0 ? total : 1));
    int64_t nact = or_merge(p, nray, offsets, hits, active);
    if (st) {
        st->merge_time += now_s() - t0;
        st->active_count += nact;
    }
    int32_t rc = 0;
    if (!c->f.trace_only) rc = or_update(c, p, active, nact, tick, tick_time, st, nthreads);
    free(active);
    free(hits);
    free(offsets);
    return rc;
}


/* ------------------------------------------------------------- emission */
/* Philox4x32-10 (Salmon et al., SC'11): ten rounds of the two multiplies
 * and the Weyl key schedule */
static void philox(uint32_t c[4], uint32_t k0, uint32_t k1) {
    for (int r = 0; r < 10; r++) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
        const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c[1] ^ k0, n1 = (uint32_t)p1;
        const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c[3] ^ k1, n3 = (uint32_t)p0;
        c[0] = n0;
        c[1] = n1;
        c[2] = n2;
        c[3] = n3;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
}
static double unif(uint32_t x) { return ((double)x + 0.5) * (1.0 / 4294967296.0); }

/* count of the n ascending values a[i] / div that are <= u */
static int64_t count_le(const double *a, int64_t n, double div, double u) {
    int64_t k = 0;
    while (k < n && a[k] / div <= u) k++;
    return k;
}

int64_t or_emit_rays(int64_t nsrc, const double *pos, const double *ngd, const int32_t *field,
                     const int32_t *spec, int64_t *lastemit, const int64_t *spec_off,
                     const double *spec_eng, const double *spec_cdf, int32_t nfield,
                     const double *boxsizes, int64_t tick, double tick_time, int64_t nray,
                     int64_t per_field, uint32_t seed, or_ray *out, int64_t cap) {
    const uint32_t K1 = 0x50535248u;
    const uint32_t tlo = (uint32_t)tick, thi = (uint32_t)((uint64_t)tick >> 32);
    int64_t *grp = malloc(sizeof(int64_t) * (nsrc > 0 ? nsrc : 1));
    double *w = malloc(sizeof(double) * (nsrc > 0 ? nsrc : 1));
    double *cdf = malloc(sizeof(double) * (nsrc > 0 ? nsrc : 1));
    uint32_t *cnt = calloc(nsrc > 0 ? nsrc : 1, sizeof(uint32_t));
    uint8_t *emitted = calloc(nsrc > 0 ? nsrc : 1, 1);
    int64_t nout = 0;
    const int32_t ngroups = per_field == 0 ? 1 : nfield;
    /* the source weights, src/sources.c:67-78 */
    for (int64_t i = 0; i < nsrc; i++) w[i] = ngd[i] * (double)(tick - lastemit[i]) * tick_time;
    for (int32_t g = 0; g < ngroups; g++) {
        int64_t ng = 0;
        for (int64_t i = 0; i < nsrc; i++)
            if (per_field == 0 || field[i] == g) grp[ng++] = i;
        const int64_t nr = per_field == 0 ? nray : per_field;
        const int64_t id0 = per_field == 0 ? 0 : (int64_t)g * per_field;
        if (nr == 0 || ng == 0) continue;
        double run = 0.0;
        for (int64_t k = 0; k < ng; k++) {
            run += w[grp[k]];
            cdf[k] = run;
        }
        if (!(cdf[ng - 1] > 0)) continue;
        if (nout + nr > cap) {
            nout = -1;
            break;
        }
        /* each ray's source (a discrete draw by weight, src/loop.c:174-190) */
        int64_t *src_of = malloc(sizeof(int64_t) * nr);
        for (int64_t k = 0; k < ng; k++) cnt[grp[k]] = 0;
        for (int64_t r = 0; r < nr; r++) {
            uint32_t c[4] = {(uint32_t)(id0 + r), tlo, thi, 0u};
            philox(c, seed, K1);
            int64_t pick = count_le(cdf, ng, cdf[ng - 1], unif(c[0]));
            if (pick > ng - 1) pick = ng - 1;
            src_of[r] = grp[pick];
            cnt[grp[pick]]++;
        }
        /* rays ordered by source, each source's in ray-id order (the
         * reference sorts by (type, isrc), src/loop.c:236) */
        int64_t *start = malloc(sizeof(int64_t) * (nsrc > 0 ? nsrc : 1));
        int64_t acc = 0;
        for (int64_t k = 0; k < ng; k++) {
            start[grp[k]] = acc;
            acc += cnt[grp[k]];
        }
        for (int64_t r = 0; r < nr; r++) {
            const int64_t s = src_of[r];
            or_ray *o = out + nout + start[s]++;
            uint32_t c[4] = {(uint32_t)(id0 + r), tlo, thi, 0u};
            philox(c, seed, K1);
            /* isotropic direction: gsl_ran_dir_3d's rejection method
             * (src/loop.c:250), two tries per block of the stream */
            double x = 0.0, y = 0.0, z = 1.0;
            int done = 0;
            for (uint32_t blk = 1; blk <= 32 && !done; blk++) {
                uint32_t d[4] = {(uint32_t)(id0 + r), tlo, thi, blk};
                philox(d, seed, K1);
                for (int h = 0; h < 2 && !done; h++) {
                    const double xx = -1.0 + 2.0 * unif(d[2 * h]);
                    const double yy = -1.0 + 2.0 * unif(d[2 * h + 1]);
                    const double ss = xx * xx + yy * yy;
                    if (ss <= 1.0) {
                        const double sc = 2.0 * sqrt(1.0 - ss);
                        x = xx * sc;
                        y = yy * sc;
                        z = -1.0 + 2.0 * ss;
                        done = 1;
                    }
                }
            }
            for (int k = 0; k < 3; k++) o->s[k] = (float)pos[3 * s + k];
            o->dir[0] = (float)x;
            o->dir[1] = (float)y;
            o->dir[2] = (float)z;
            /* a source's rays share its weight, src/loop.c:305-321 */
            o->Nph = w[s] * (1.0 / (double)cnt[s]);
            /* photon energy: a bin of the source's spectrum, src/spectra.c:50-60 */
            const int64_t a = spec_off[spec[s]], nb = spec_off[spec[s] + 1] - a;
            int64_t ib = 0;
            const double u3 = unif(c[3]);
            while (ib < nb && spec_cdf[a + ib] <= u3) ib++;
            if (ib > nb - 1) ib = nb - 1;
            o->freq = spec_eng[a + ib];
            o->length = boxsizes[field[s]] * 2;  /* src/loop.c:267 */
            o->type = -1;
            o->field = field[s];
            o->src_or_par = s;
            emitted[s] = 1;
        }
        nout += nr;
        free(start);
        free(src_of);
    }
    if (nout >= 0)
        for (int64_t i = 0; i < nsrc; i++)
            if (emitted[i]) lastemit[i] = tick;  /* src/loop.c:335-338 */
    free(grp);
    free(w);
    free(cdf);
    free(cnt);
    free(emitted);
    return nout;
}


/* ------------------------------------------------- recombination rays */
void or_lte_tables(const or_ctx *c, double *eng, double *cdf, double *tb) {
    const or_units *u = &c->u;
    const double th[3] = {u->C_HI_ENERGY, u->C_HEI_ENERGY, u->C_HEII_ENERGY};
    const double logT_step = (10.0 - 3.0) / 128;
    for (int s = 0; s < 3; s++) {
        const double freq_step = (16 * th[s] - th[s]) / 1024;
        for (int j = 0; j < 1024; j++) eng[s * 1024 + j] = th[s] + freq_step * j;
        for (int i = 0; i < 128; i++) {
            const double kT = u->C_BOLTZMANN * exp10(3.0 + logT_step * i);
            double *row = cdf + ((int64_t)s * 128 + i) * 1024;
            double sum = 0.0;
            for (int j = 0; j < 1024; j++) {
                const double freq = eng[s * 1024 + j];
                sum += or_xs_get(c, s, freq) * freq * freq * freq * exp(-(freq - th[s]) / kT);
                row[j] = sum;
            }
            for (int j = 0; j < 1024; j++) row[j] = sum > 0 ? row[j] / sum : 1.0;
        }
    }
    for (int k = 0; k < 127; k++) tb[k] = exp10(3.0 + logT_step * (k + 1));
}

typedef struct {
    int64_t ipar, k;
} rec_draw;
static int rec_draw_cmp(const void *a, const void *b) {
    const rec_draw *x = a, *y = b;
    if (x->ipar != y->ipar) return x->ipar < y->ipar ? -1 : 1;
    return (x->k > y->k) - (x->k < y->k);
}

int64_t or_emit_recomb(const or_ctx *c, const or_tree *t, or_psys *p, const or_ray *prev, int64_t nprev,
                       int64_t cut, int64_t nrec, double threshold, int32_t disable_lte,
                       const double *lte_eng, const double *lte_cdf, const double *lte_tb, double boxsize,
                       int64_t tick, uint32_t seed, or_ray *out, int64_t cap) {
    if (c->f.on_the_spot || nrec == 0 || tick <= 1) return 0;
    const uint32_t K1 = 0x50535248u;
    const uint32_t tlo = (uint32_t)tick, thi = (uint32_t)((uint64_t)tick >> 32);
    const int64_t np = p->npar;
    /* merge_pars' active set up to x_src_length: every particle the first
     * nr rays hit (src/loop.c:620-637) */
    uint8_t *hit = calloc(np > 0 ? np : 1, 1);
    const int64_t nr = nprev < cut ? nprev : cut;
    if (nr > 0) {
        int64_t *offs = malloc(sizeof(int64_t) * (nr + 1));
        or_hit *h = trace_all(t, p, prev, nr, offs, 1);
        for (int64_t k = 0; k < offs[nr]; k++) hit[h[k].ipar] = 1;
        free(h);
        free(offs);
    }
    int64_t *cand = malloc(sizeof(int64_t) * (np > 0 ? np : 1));
    int64_t ncand = 0;
    for (int64_t i = 0; i < np; i++)
        if (hit[i]) cand[ncand++] = i;
    free(hit);
    float *yg[3] = {p->yGrecHII, p->yGrecHeII, p->yGrecHeIII};
    const double per_mass[3] = {c->u.C_H_PER_MASS, c->u.C_HE_PER_MASS, c->u.C_HE_PER_MASS};
    double *cum = malloc(sizeof(double) * (ncand > 0 ? ncand : 1));
    rec_draw *dr = malloc(sizeof(rec_draw) * ((nrec > ncand ? nrec : ncand) + 1));
    int64_t nout = 0;
    for (int s = 0; s < (c->f.h_only ? 1 : 3); s++) {
        int64_t nd = 0;
        if (nrec > 0) {
            if (ncand == 0) continue;
            /* weights yGrec_s / x_HI (src/loop.c:206-211), summed in chunks */
            double before = 0.0;
            for (int64_t j0 = 0; j0 < ncand; j0 += 1024) {
                const int64_t j1 = j0 + 1024 < ncand ? j0 + 1024 : ncand;
                double in_chunk = 0.0;
                for (int64_t j = j0; j < j1; j++) {
                    double x = p->lambdaH[cand[j]];
                    if (x == 0) x = 1e-10;
                    in_chunk += (double)yg[s][cand[j]] / x;
                    cum[j] = in_chunk;
                }
                for (int64_t j = j0; j < j1; j++) cum[j] = before + cum[j];
                before += in_chunk;
            }
            const double total = before;
            if (!(total > 0)) continue;
            for (int64_t k = 0; k < nrec; k++) {
                uint32_t ctr[4] = {(uint32_t)k, tlo, thi, (uint32_t)(s + 1) << 16};
                philox(ctr, seed, K1);
                const double u = unif(ctr[0]);
                int64_t lo = 0, hi = ncand;
                while (lo < hi) {
                    const int64_t mid = (lo + hi) / 2;
                    if (cum[mid] / total <= u) lo = mid + 1;
                    else hi = mid;
                }
                dr[nd].ipar = cand[lo < ncand - 1 ? lo : ncand - 1];
                dr[nd].k = k;
                nd++;
            }
        } else {
            /* one ray per candidate above the threshold (src/loop.c:222-229) */
            for (int64_t j = 0; j < ncand; j++)
                if ((double)yg[s][cand[j]] > threshold) {
                    dr[nd].ipar = cand[j];
                    dr[nd].k = nd;
                    nd++;
                }
        }
        if (nd == 0) continue;
        if (nout + nd > cap) {
            nout = -1;
            break;
        }
        /* ordered by (type, particle), src/loop.c:236 */
        qsort(dr, (size_t)nd, sizeof(rec_draw), rec_draw_cmp);
        for (int64_t a = 0; a < nd;) {
            int64_t b = a;
            while (b < nd && dr[b].ipar == dr[a].ipar) b++;
            const int64_t ip = dr[a].ipar;
            const double N = (double)p->mass[ip] * per_mass[s];
            for (int64_t q = a; q < b; q++) {
                or_ray *o = out + nout + q;
                for (int d = 0; d < 3; d++) o->s[d] = p->pos[3 * ip + d];
                /* src/loop.c:283-293 and the shared weight of :305-321 */
                o->Nph = ((double)yg[s][ip] * N) * (1.0 / (double)(b - a));
                uint32_t c0[4] = {(uint32_t)dr[q].k, tlo, thi, (uint32_t)(s + 1) << 16};
                philox(c0, seed, K1);
                if (disable_lte) {
                    o->freq = lte_eng[s * 1024];
                } else {
                    /* lte_gen_freq, src/lte.c:34-45 */
                    const double T = or_ieye2T(c, p->ie[ip], psys_ye(c, p, ip));
                    int row = 0;
                    while (row < 127 && lte_tb[row] <= T) row++;
                    const double *cd = lte_cdf + ((int64_t)s * 128 + row) * 1024;
                    const double u3 = unif(c0[3]);
                    int bin = 0;
                    while (bin < 1024 && cd[bin] <= u3) bin++;
                    o->freq = lte_eng[s * 1024 + (bin < 1023 ? bin : 1023)];
                }
                double x = 0.0, y = 0.0, z = 1.0;
                int done = 0;
                for (uint32_t blk = 1; blk <= 32 && !done; blk++) {
                    uint32_t d4[4] = {(uint32_t)dr[q].k, tlo, thi, ((uint32_t)(s + 1) << 16) | blk};
                    philox(d4, seed, K1);
                    for (int h = 0; h < 2 && !done; h++) {
                        const double xx = -1.0 + 2.0 * unif(d4[2 * h]);
                        const double yy = -1.0 + 2.0 * unif(d4[2 * h + 1]);
                        const double ss = xx * xx + yy * yy;
                        if (ss <= 1.0) {
                            const double sc = 2.0 * sqrt(1.0 - ss);
                            x = xx * sc;
                            y = yy * sc;
                            z = -1.0 + 2.0 * ss;
                            done = 1;
                        }
                    }
                }
                o->dir[0] = (float)x;
                o->dir[1] = (float)y;
                o->dir[2] = (float)z;
                o->length = boxsize * 2 * 0.1; /* src/loop.c:169 */
                o->type = s;
                o->field = 0;
                o->src_or_par = ip;
            }
            /* the emitted photons leave the recombination count
             * (src/loop.c:339-368) */
            for (int64_t q = a; q < b; q++) {
                yg[s][ip] = (float)((double)yg[s][ip] - out[nout + q].Nph / N);
                if (yg[s][ip] < 0) yg[s][ip] = 0;
            }
            a = b;
        }
        nout += nd;
    }
    free(cand);
    free(cum);
    free(dr);
    return nout;
}
