/*
 * psr_host.c -- host-side set-up that keeps the reference's formats:
 * internal units (src/units.c), the atomic-rate table file
 * (src/tabfun.c:60-130, src/atomic-rate.c:253-278), the analytic
 * photo-ionization cross sections (src/atomic-rate.c:280-318 with the fits
 * of src/verner.F90), the SPH column kernel table (src/sph-depth.c) and the
 * secondary-ionization table (src/secion.c).  The device kernels read the
 * tables built here.
 */
#define _GNU_SOURCE
#include <math.h>
#include <stddef.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <strings.h>
#include <sys/types.h>

#include "psphray_amd.h"

int psr_units_init(psr_units *u, double h, double hmf, double hemf, double cmh, double gramh,
                   double sech) {
    if (!u || h <= 0 || cmh <= 0 || gramh <= 0 || sech <= 0) return PSR_ERR_ARG;
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
    return PSR_OK;
}

/* ---------------------------------------------------------------- rates */
/* Reads the fixed 14-column Fortran layout: header "min max nrows step",
 * a line of column names, one ignored citation line, then nrows rows.
 * Lines starting with '#' are comments. */
typedef struct {
    int nrows, ncols;
    double min, step;
    double **col;
    char **name;
} rate_file;

static void rate_file_free(rate_file *r) {
    for (int i = 0; i < r->ncols; i++) {
        free(r->col[i]);
        free(r->name[i]);
    }
    free(r->col);
    free(r->name);
}

static int rate_file_read(rate_file *r, const char *path, char *err, size_t errlen) {
    memset(r, 0, sizeof(*r));
    FILE *fp = fopen(path, "r");
    if (!fp) {
        snprintf(err, errlen, "cannot open rates file %s", path);
        return -1;
    }
    char *line = NULL;
    size_t cap = 0;
    ssize_t len;
    int stage = 0, row = 0;
    while ((len = getline(&line, &cap, fp)) >= 0) {
        if (line[0] == '#') continue;
        if (stage == 0) {
            double mx;
            if (sscanf(line, "%lf %lf %d %lf", &r->min, &mx, &r->nrows, &r->step) != 4 ||
                r->nrows < 2) {
                snprintf(err, errlen, "%s: expecting 'min max nrows step'", path);
                break;
            }
            stage = 1;
        } else if (stage == 1) {
            r->ncols = (int)(len / 14);
            r->col = calloc(r->ncols, sizeof(double *));
            r->name = calloc(r->ncols, sizeof(char *));
            for (int i = 0; i < r->ncols; i++) {
                char buf[15];
                memcpy(buf, line + 14 * i, 14);
                buf[14] = 0;
                char *b = buf, *e = buf + 13;
                while (e >= b && (*e == ' ' || *e == '\n' || *e == 0)) *e-- = 0;
                while (*b == ' ') b++;
                r->name[i] = strdup(b);
                r->col[i] = calloc(r->nrows, sizeof(double));
            }
            stage = 2;
        } else if (stage == 2) {
            stage = 3; /* citation line */
        } else {
            for (int i = 0; i < r->ncols; i++) {
                if (14 * i > len) {
                    snprintf(err, errlen, "%s: short row %d", path, row);
                    goto done;
                }
                char buf[14];
                memcpy(buf, line + 14 * i, 13);
                buf[13] = 0;
                r->col[i][row] = atof(buf);
            }
            if (++row == r->nrows) {
                stage = 4;
                break;
            }
        }
    }
done:
    free(line);
    fclose(fp);
    if (stage != 4) {
        if (!err[0]) snprintf(err, errlen, "%s: table incomplete (%d rows)", path, row);
        rate_file_free(r);
        return -1;
    }
    return 0;
}

static int rate_col(const rate_file *r, const char *name) {
    for (int i = 0; i < r->ncols; i++)
        if (!strcasecmp(r->name[i], name)) return i;
    return -1;
}

/* ------------------------------------------------------- cross sections */
/* Verner et al. (1996) H I fit and Osterbrock's He I / He II forms, in the
 * single-precision arithmetic of src/verner.F90:5-65 */
static float xs_verner_hi(float Ry) {
    const float E0 = (float)4.298e-1, sig0 = (float)5.475e4, ya = (float)3.288e1,
                P = (float)2.963;
    if (Ry < 0.5f) return 0.0f;
    const float eV = (float)((double)Ry * 13.6);
    const float x = eV / E0;
    const float xm1 = x - 1.0f;
    const float a = sig0 * (xm1 * xm1);
    const double b = pow((double)x, 0.5 * (double)P - 5.5);
    const float c = powf(1.0f + sqrtf(x / ya), -P);
    float sigma = (float)((double)a * b * (double)c);
    sigma = (float)((double)sigma * 1.0e-18);
    return sigma < 0 ? 0.0f : sigma;
}
static float xs_hei(float Ry) {
    const float nu = 24.587f;
    if (2.0f * Ry * 13.6f < nu) return 0.0f;
    const float sf = Ry * 13.6f / nu;
    return 7.2e-18f * (1.66f * powf(sf, -2.05f) + 0.66f * powf(sf, -3.05f));
}
static float xs_heii(float Ry) {
    const float nu = 54.416f;
    if (2.0f * Ry * 13.6f < nu) return 0.0f;
    const float t = Ry * 13.6f / nu;
    return 1.58e-18f * (1.0f / (t * t * t));
}

/* ------------------------------------------------------------ SPH kernel */
/* cubic spline W h^3 = 8/pi {1 - 6r^2 + 6r^3, r <= 1/2; 2(1-r)^3, r <= 1}
 * (src/sph-depth.c:208-221) */
static double sph_wh3(double r) {
    const double norm = 8.0 / M_PI;
    if (r <= 0.5) return norm * (1.0 - 6.0 * r * r + 6.0 * r * r * r);
    if (r <= 1.0) {
        const double d = 1.0 - r;
        return norm * 2.0 * d * d * d;
    }
    return 0.0;
}

/* 20-point Gauss-Legendre nodes/weights on [-1, 1] (positive half) */
static const double GL_X[10] = {0.0765265211334973, 0.2277858511416451, 0.3737060887154195,
                                0.5108670019508271, 0.6360536807265150, 0.7463319064601508,
                                0.8391169718222188, 0.9122344282513259, 0.9639719272779138,
                                0.9931285991850949};
static const double GL_W[10] = {0.1527533871307258, 0.1491729864726037, 0.1420961093183820,
                                0.1316886384491766, 0.1181945319615184, 0.1019301198172404,
                                0.0832767415767048, 0.0626720483341091, 0.0406014298003869,
                                0.0176140071391521};

static double integrate_w(double b, double z0, double z1, int panels) {
    double sum = 0.0;
    const double hw = (z1 - z0) / panels;
    for (int p = 0; p < panels; p++) {
        const double a = z0 + p * hw, m = a + 0.5 * hw, h2 = 0.5 * hw;
        for (int k = 0; k < 10; k++) {
            const double za = m - h2 * GL_X[k], zb = m + h2 * GL_X[k];
            sum += GL_W[k] * h2 * (sph_wh3(sqrt(b * b + za * za)) + sph_wh3(sqrt(b * b + zb * zb)));
        }
    }
    return sum;
}

/* column integral of W along a line at impact parameter b (in units of h),
 * split where r crosses 1/2 so each piece is smooth */
static double sph_column(double b) {
    if (b >= 1.0) return 0.0;
    const double zmax = sqrt(1.0 - b * b);
    double s = 0.0;
    if (b < 0.5) {
        const double zk = sqrt(0.25 - b * b);
        s = integrate_w(b, 0.0, zk, 64) + integrate_w(b, zk, zmax, 64);
    } else {
        s = integrate_w(b, 0.0, zmax, 64);
    }
    return 2.0 * s;
}

/* --------------------------------------------------- secondary ionization */
/* Ricotti, Gnedin & Shull (2002) fits as tabulated by src/secion.c:21-73,
 * including its threshold "gate - 0.1" in internal energy units */
static const double SI_A[3] = {0.2, 0.2, 0.4};
static const double SI_B[3][2] = {{0.4092, 0.38}, {0.4614, 0.38}, {0.2663, 0.34}};
static const double SI_CC[3][2] = {{1.7592, 2.0}, {1.6660, 2.0}, {1.3163, 2.0}};
static const double SI_K[3][2] = {{0.3908, 0.6941}, {0.0554, 0.0984}, {1.0, 3.9811}};

static double si_phi(double uev, int j, double E, double x) {
    const double gate = 28 * uev;
    const int f1 = (E >= (gate - 0.1));
    const double f2 = f1 * pow(gate / E, 0.4);
    const double y1 = SI_K[j][0] * pow(1. - pow(x, SI_B[j][0]), SI_CC[j][0]);
    const double y2 = SI_K[j][1] * pow(x, SI_A[j]) * pow(1 - pow(x, SI_B[j][1]), SI_CC[j][1]);
    return (y1 * f1 - y2 * f2);
}
static double si_eh(double uev, double E, double x) {
    const double gate = 11 * uev;
    const int f1 = (E >= gate - 0.1);
    const double f2 = f1 * pow(gate / E, 0.7);
    const double y1 = SI_K[2][0] * pow(1 - pow(x, SI_B[2][0]), SI_CC[2][0]);
    const double y2 = SI_K[2][1] * pow(x, SI_A[2]) * pow(1 - pow(x, SI_B[2][1]), SI_CC[2][1]);
    return (1. - y1 * f1 + y2 * f2);
}

/* ----------------------------------------------------------------- build */
int psr_tables_build(psr_tables *t, const psr_units *u, const char *rates_path, int with_secion,
                     char *err, size_t errlen) {
    char dummy[8];
    if (!err) {
        err = dummy;
        errlen = sizeof(dummy);
    }
    err[0] = 0;
    if (!t || !u || !rates_path) return PSR_ERR_ARG;
    memset(t, 0, sizeof(*t));
    rate_file rf;
    if (rate_file_read(&rf, rates_path, err, errlen)) return PSR_ERR_ARG;

    /* columns and their units, src/atomic-rate.c:257-271 */
    static const char *names[15] = {"HIci", "HIIrcA", "HIIrccA", "HIIrcB", "HIIrccB",
                                    "HeIci", "HeIIci", "HeIIrcA", "HeIIrcB", "HeIIIrcA",
                                    "HeIIIrcB", "HIcic", "HIcec", "Brem", "Compton"};
    const double cm3s = u->U_CM * u->U_CM * u->U_CM / u->U_SEC;
    const double ecm3s = u->U_ERG * u->U_CM * u->U_CM * u->U_CM / u->U_SEC;
    const double units[15] = {cm3s, cm3s, ecm3s, cm3s, ecm3s, cm3s, cm3s, cm3s,
                              cm3s, cm3s, cm3s, ecm3s, ecm3s, ecm3s, u->U_ERG / u->U_SEC};
    const int n = rf.nrows;
    double *ar = malloc(sizeof(double) * 16 * (size_t)n);
    memcpy(ar, rf.col[0], sizeof(double) * n);
    for (int k = 0; k < 15; k++) {
        double *dst = ar + (size_t)(k + 1) * n;
        const int c = rate_col(&rf, names[k]);
        if (c >= 0) {
            for (int j = 0; j < n; j++) dst[j] = rf.col[c][j] * units[k];
        } else if (k == 13) {
            /* analytic Bremsstrahlung (Cen 1992) evaluated on column 0 as
             * tabfun_setup_col does (src/tabfun.c:32-34) */
            for (int j = 0; j < n; j++) dst[j] = 1.42e-27 * 1.5 * sqrt(rf.col[0][j]) * units[k];
        } else if (k == 14) {
            for (int j = 0; j < n; j++) dst[j] = 0.0 * units[k];
        } else {
            snprintf(err, errlen, "rates file %s lacks column %s", rates_path, names[k]);
            free(ar);
            rate_file_free(&rf);
            return PSR_ERR_ARG;
        }
    }
    t->ar_nrows = n;
    t->ar_min = rf.min;
    t->ar_step_inv = 1.0 / rf.step;
    t->ar = ar;
    rate_file_free(&rf);

    /* cross sections, src/atomic-rate.c:284-309 */
    const int nx = 16384;
    double *xs = malloc(sizeof(double) * 4 * (size_t)nx);
    const double xmin = 1 * u->C_HI_ENERGY, xmax = 100 * u->C_HI_ENERGY;
    const double xstep = (xmax - xmin) / (nx - 1);
    const double cm2 = u->U_CM * u->U_CM;
    for (int i = 0; i < nx; i++) {
        const double e = xmin + xstep * i;
        const float ryd = (float)(e / u->C_HI_ENERGY);
        xs[i] = e;
        xs[nx + i] = (double)xs_verner_hi(ryd) * cm2;
        xs[2 * nx + i] = (double)xs_hei(ryd) * cm2;
        xs[3 * nx + i] = (double)xs_heii(ryd) * cm2;
    }
    t->xs_nrows = nx;
    t->xs_min = xmin;
    t->xs_step_inv = 1.0 / xstep;
    t->xs_cut[0] = 13.6 * u->U_EV;
    t->xs_cut[1] = 24.5 * u->U_EV;
    t->xs_cut[2] = 54.4 * u->U_EV;
    t->xs = xs;

    /* SPH column kernel: 4096 float32-rounded samples at b = 4n/16383, the
     * sampling of the reference table src/kernel.akline */
    const int nk = 4096;
    double *kt = malloc(sizeof(double) * nk);
    for (int i = 0; i < nk; i++) kt[i] = (double)(float)sph_column(4.0 * i / 16383.0);
    t->kernel_n = nk;
    t->kernel = kt;

    if (with_secion) {
        /* src/secion.c:75-99 */
        const int n1 = 4096, n2 = 1024;
        const double min1 = 10 * u->U_EV, max1 = 200 * u->U_EV, min2 = -8, max2 = 0;
        const double step1 = (max1 - min1) / n1, step2 = (max2 - min2) / n2;
        double *sx = malloc(sizeof(double) * n1), *sy = malloc(sizeof(double) * n2);
        for (int i = 0; i < n1; i++) sx[i] = min1 + i * step1;
        for (int j = 0; j < n2; j++) sy[j] = min2 + j * step2;
        double *phi = malloc(sizeof(double) * (size_t)n1 * n2);
        double *phe = malloc(sizeof(double) * (size_t)n1 * n2);
        double *eh = malloc(sizeof(double) * (size_t)n1 * n2);
        for (int i = 0; i < n1; i++)
            for (int j = 0; j < n2; j++) {
                const double x = pow(10, sy[j]);
                const size_t o = (size_t)i * n2 + j;
                phi[o] = si_phi(u->U_EV, 0, sx[i], x);
                phe[o] = si_phi(u->U_EV, 1, sx[i], x);
                eh[o] = si_eh(u->U_EV, sx[i], x);
            }
        t->si_n1 = n1;
        t->si_n2 = n2;
        t->si_min1 = min1;
        t->si_min2 = min2;
        t->si_step1_inv = 1.0 / step1;
        t->si_step2_inv = 1.0 / step2;
        t->si_x = sx;
        t->si_y = sy;
        t->si_phi_hi = phi;
        t->si_phi_hei = phe;
        t->si_eh = eh;
    }
    return PSR_OK;
}

/* the SPH column table as data: src/kernel.akline, one value per line,
 * each followed by a comma (the file is #included as an array initialiser
 * at src/sph-depth.c:3-6).  Replaces the quadrature table of
 * psr_tables_build; the entry count is the file's (sph_depth indexes it with
 * n = (int)(N r_h), src/sph-depth.c:26-33). */
int psr_tables_load_kernel(psr_tables *t, const char *path, char *err, size_t errlen) {
    if (!t || !path) return PSR_ERR_ARG;
    FILE *f = fopen(path, "r");
    if (!f) {
        if (err && errlen) snprintf(err, errlen, "cannot open %s", path);
        return PSR_ERR_ARG;
    }
    size_t cap = 4096, n = 0;
    double *v = malloc(sizeof(double) * cap);
    char line[256];
    int bad = 0;
    while (fgets(line, sizeof line, f)) {
        char *p = line;
        while (*p == ' ' || *p == '\t') p++;
        if (*p == '\n' || *p == '\r' || *p == 0) continue;
        char *end = NULL;
        const double x = strtod(p, &end);
        if (end == p) { bad = 1; break; }
        while (*end == ' ' || *end == '\t') end++;
        if (*end != ',' && *end != '\n' && *end != '\r' && *end != 0) { bad = 1; break; }
        if (n == cap) {
            cap *= 2;
            v = realloc(v, sizeof(double) * cap);
        }
        v[n++] = x;
    }
    fclose(f);
    if (bad || n == 0 || n > (1u << 24)) {
        free(v);
        if (err && errlen) snprintf(err, errlen, "%s: not a kernel table (one number per line)", path);
        return PSR_ERR_ARG;
    }
    free((void *)t->kernel);
    t->kernel_n = (int32_t)n;
    t->kernel = v;
    return PSR_OK;
}

void psr_tables_free(psr_tables *t) {
    if (!t) return;
    free((void *)t->ar);
    free((void *)t->xs);
    free((void *)t->kernel);
    free((void *)t->si_x);
    free((void *)t->si_y);
    free((void *)t->si_phi_hi);
    free((void *)t->si_phi_hei);
    free((void *)t->si_eh);
    memset(t, 0, sizeof(*t));
}

/* clamped linear interpolation, src/tabfun.c:38-58 */
static double lin_get(const double *xcol, const double *ycol, int nrows, double min,
                      double step_inv, double v) {
    const ssize_t idx = (ssize_t)((v - min) * step_inv);
    if (idx < 0) return ycol[0];
    if (idx >= nrows - 1) return ycol[nrows - 1];
    const double left = xcol[idx], right = xcol[idx + 1];
    return ((right - v) * ycol[idx] + (v - left) * ycol[idx + 1]) * step_inv;
}

double psr_tab_ar_get(const psr_tables *t, int col, double logT) {
    const int n = t->ar_nrows;
    return lin_get(t->ar, t->ar + (size_t)(col + 1) * n, n, t->ar_min, t->ar_step_inv, logT);
}

double psr_tab_xs_get(const psr_tables *t, int id, double eng) {
    if (eng < t->xs_cut[id]) return 0.0;
    const int n = t->xs_nrows;
    return lin_get(t->xs, t->xs + (size_t)(id + 1) * n, n, t->xs_min, t->xs_step_inv, eng);
}

double psr_tab_sph_depth(const psr_tables *t, double r_h) {
    const int n = (int)(t->kernel_n * r_h);
    return (n < t->kernel_n && n >= 0) ? t->kernel[n] : 0.0;
}

/* ------------------------------------------------------------ LTE spectra */
/* lte_init / lte_fill (src/lte.c:28-72): per species the photon energies
 * E_j = E0 + j (16 E0 - E0) / 1024 and, per row i of log10 T = 3 + i 7/128,
 * the weights sigma(E_j) E_j^3 exp(-(E_j - E0) / kT) (in that order of
 * operations), here as cumulative sums divided by the row's total; the
 * rows' lower bounds in T, tbound[k] = 10^(3 + (k + 1) 7/128), pick a row
 * for a temperature without a logarithm */
int psr_lte_tables(const psr_tables *t, const psr_units *u, double *energy, double *cdf, double *tbound) {
    if (!t || !u || !energy || !cdf || !tbound || !t->xs) return PSR_ERR_ARG;
    const double e0[3] = {u->C_HI_ENERGY, u->C_HEI_ENERGY, u->C_HEII_ENERGY};
    const double step = (PSR_LTE_LOGT_MAX - PSR_LTE_LOGT_MIN) / PSR_LTE_NLOGT;
    for (int s = 0; s < 3; s++) {
        const double fstep = (16 * e0[s] - e0[s]) / PSR_LTE_NFREQ;
        double *eng = energy + (size_t)s * PSR_LTE_NFREQ;
        for (int j = 0; j < PSR_LTE_NFREQ; j++) eng[j] = e0[s] + fstep * j;
        for (int i = 0; i < PSR_LTE_NLOGT; i++) {
            const double kT = u->C_BOLTZMANN * exp10(PSR_LTE_LOGT_MIN + step * i);
            double *row = cdf + ((size_t)s * PSR_LTE_NLOGT + i) * PSR_LTE_NFREQ;
            double run = 0.0;
            for (int j = 0; j < PSR_LTE_NFREQ; j++) {
                const double f = eng[j];
                run += psr_tab_xs_get(t, s, f) * f * f * f * exp(-(f - e0[s]) / kT);
                row[j] = run;
            }
            for (int j = 0; j < PSR_LTE_NFREQ; j++) row[j] = run > 0 ? row[j] / run : 1.0;
        }
    }
    for (int k = 0; k < PSR_LTE_NLOGT - 1; k++) tbound[k] = exp10(PSR_LTE_LOGT_MIN + step * (k + 1));
    return PSR_OK;
}

/* ------------------------------------------------------------ snapshots */
/* Gadget-style snapshots: a 256-byte header, then the format's blocks in a
 * fixed order, each as <int32 size> data <int32 size>, present when its
 * length is non-zero (_offset, src/reader-psphray.inc:332-344).  Two
 * formats, as reader_new names them (src/reader.c:36-46):
 *   psphray       src/reader-psphray.inc: order :55-66, item sizes :68-90,
 *                 lengths :115-330 (gas type 0, stars 4, sources 5);
 *   massiveblack  src/reader-massiveblack.inc ("e5" too): order, item sizes
 *                 and lengths :26-330 (also type 1, dark matter).
 * The lengths read N[] and the sfr/sft/met/Hmf/.. flags, all in the
 * header prefix the two formats share. */
typedef struct {
    const char *name;
    int64_t item;
    /* length = sum of N[t] over the type mask bits, when the flag (a byte
     * offset into the header, or -1) is set */
    unsigned types;
    int flag_off;
} snap_block;

#define T0 1u
#define T1 2u
#define T4 16u
#define T5 32u
#define HOFF(f) ((int)offsetof(psr_snap_header, f))

static const snap_block SNAP_PSPHRAY[] = {
    {"header", 256, 0, -1}, {"pos", 12, T0 | T4 | T5, -1}, {"id", 8, T0 | T4 | T5, -1},
    {"mass", 4, T0 | T4 | T5, -1}, {"ie", 4, T0, -1}, {"rho", 4, T0, -1}, {"ye", 4, T0, -1},
    {"xHI", 4, T0, -1}, {"sml", 4, T0, -1}, {"Hmf", 4, T0, HOFF(flag_Hmf)},
    {"Hemf", 4, T0, HOFF(flag_Hemf)}, {"xHeI", 4, T0, HOFF(flag_Helium)},
    {"xHeII", 4, T0, HOFF(flag_Helium)}, {"gammaHI", 4, T0, HOFF(flag_gammaHI)},
    {"HIa_cloudy", 4, T0, HOFF(flag_cloudy)}, {"eos", 4, T0, HOFF(flag_eos)}, {"lasthit", 8, T0, -1},
    {"hits", 4, T0, -1}, {"sft", 4, T4, -1}, {"ngammas", 8, T5, -1}, {"spec", 8, T5, -1},
};
static const snap_block SNAP_MASSIVEBLACK[] = {
    {"header", 256, 0, -1}, {"pos", 12, T0 | T1 | T4 | T5, -1}, {"vel", 12, T0 | T1 | T4 | T5, -1},
    {"id", 8, T0 | T1 | T4 | T5, -1}, {"mass", 4, T0 | T4 | T5, -1}, {"ie", 4, T0, -1},
    {"rho", 4, T0, -1}, {"ye", 4, T0, -1}, {"xHI", 4, T0, -1}, {"sml", 4, T0, -1},
    {"sfr", 4, T0, HOFF(flag_sfr)}, {"sft", 4, T4, HOFF(flag_sft)}, {"met", 4, T0 | T4, HOFF(flag_met)},
    {"bhmass", 4, T5, -1}, {"bhmdot", 4, T5, -1},
};

typedef struct {
    const snap_block *b;
    int n;
} snap_format;

static int snap_format_of(const char *format, snap_format *f) {
    if (!format || !strcmp(format, "psphray")) {
        f->b = SNAP_PSPHRAY;
        f->n = (int)(sizeof SNAP_PSPHRAY / sizeof SNAP_PSPHRAY[0]);
        return 0;
    }
    if (!strcmp(format, "massiveblack") || !strcmp(format, "e5")) {
        f->b = SNAP_MASSIVEBLACK;
        f->n = (int)(sizeof SNAP_MASSIVEBLACK / sizeof SNAP_MASSIVEBLACK[0]);
        return 0;
    }
    return -1;
}

static int snap_block_id(const snap_format *f, const char *b) {
    for (int i = 0; i < f->n; i++)
        if (!strcmp(b, f->b[i].name)) return i;
    return -1;
}

static int64_t snap_length(const snap_format *f, const psr_snap_header *h, int id) {
    const snap_block *b = &f->b[id];
    if (id == 0) return 1;
    if (b->flag_off >= 0) {
        int32_t flag;
        memcpy(&flag, (const char *)h + b->flag_off, sizeof flag);
        if (!flag) return 0;
    }
    int64_t n = 0;
    for (int t = 0; t < 6; t++)
        if (b->types & (1u << t)) n += h->N[t];
    return n;
}

int psr_snapshot_block_info(const char *format, const psr_snap_header *h, const char *block, int64_t *length,
                            int64_t *itemsize, int64_t *offset) {
    snap_format f;
    if (!h || !block || snap_format_of(format, &f)) return PSR_ERR_ARG;
    const int id = snap_block_id(&f, block);
    if (id < 0) return PSR_ERR_ARG;
    int64_t off = 0;
    for (int i = 0; i < id; i++) {
        const int64_t bytes = snap_length(&f, h, i) * f.b[i].item;
        if (bytes > 0) off += bytes + 8;
    }
    if (length) *length = snap_length(&f, h, id);
    if (itemsize) *itemsize = f.b[id].item;
    if (offset) *offset = off;
    return PSR_OK;
}

static int snap_err(char *err, size_t errlen, const char *msg, const char *what) {
    if (err && errlen) snprintf(err, errlen, "%s: %s", what, msg);
    return PSR_ERR_ARG;
}

int psr_snapshot_read_header(const char *path, psr_snap_header *h, char *err, size_t errlen) {
    if (!path || !h) return PSR_ERR_ARG;
    FILE *f = fopen(path, "rb");
    if (!f) return snap_err(err, errlen, "cannot open", path);
    int32_t m0 = 0, m1 = 0;
    memset(h, 0, sizeof(*h));
    const int ok = fread(&m0, 4, 1, f) == 1 && fread(h, 256, 1, f) == 1 && fread(&m1, 4, 1, f) == 1;
    fclose(f);
    if (!ok || m0 != 256 || m1 != 256) return snap_err(err, errlen, "not a psphray snapshot (header record)", path);
    return PSR_OK;
}

int psr_snapshot_read_block(const char *format, const char *path, const psr_snap_header *h, const char *block,
                            void *buf, char *err, size_t errlen) {
    int64_t len, item, off;
    if (!path || !h || !block || psr_snapshot_block_info(format, h, block, &len, &item, &off))
        return snap_err(err, errlen, "unknown block or format", block ? block : "(null)");
    if (len == 0) return PSR_OK;
    if (!buf) return PSR_ERR_ARG;
    FILE *f = fopen(path, "rb");
    if (!f) return snap_err(err, errlen, "cannot open", path);
    const int64_t bytes = len * item;
    int32_t m0 = 0, m1 = 0;
    int ok = fseeko(f, off, SEEK_SET) == 0 && fread(&m0, 4, 1, f) == 1 && m0 == (int32_t)bytes &&
             fread(buf, 1, (size_t)bytes, f) == (size_t)bytes && fread(&m1, 4, 1, f) == 1 && m1 == (int32_t)bytes;
    fclose(f);
    if (!ok) return snap_err(err, errlen, "block size markers do not match the header", block);
    return PSR_OK;
}

int psr_snapshot_write(const char *format, const char *path, const psr_snap_header *h, const char *const *names,
                       const void *const *bufs, int32_t nblocks, char *err, size_t errlen) {
    snap_format fm;
    if (!path || !h || (nblocks > 0 && (!names || !bufs))) return PSR_ERR_ARG;
    if (snap_format_of(format, &fm)) return snap_err(err, errlen, "unknown format", format);
    for (int32_t k = 0; k < nblocks; k++)
        if (!names[k] || snap_block_id(&fm, names[k]) <= 0)
            return snap_err(err, errlen, "unknown block", names[k] ? names[k] : "(null)");
    for (int id = 1; id < fm.n; id++)
        if (snap_length(&fm, h, id) * fm.b[id].item > INT32_MAX)
            return snap_err(err, errlen, "block over 2 GiB (its int32 size marker)", fm.b[id].name);
    FILE *f = fopen(path, "wb");
    if (!f) return snap_err(err, errlen, "cannot create", path);
    int ok = 1;
    for (int id = 0; id < fm.n && ok; id++) {
        const int64_t bytes = snap_length(&fm, h, id) * fm.b[id].item;
        if (bytes == 0) continue;
        const void *src = id == 0 ? (const void *)h : NULL;
        for (int32_t k = 0; k < nblocks && id > 0; k++)
            if (!strcmp(names[k], fm.b[id].name)) src = bufs[k];
        const int32_t m = (int32_t)bytes;
        ok = fwrite(&m, 4, 1, f) == 1;
        if (src) {
            ok = ok && fwrite(src, 1, (size_t)bytes, f) == (size_t)bytes;
        } else {
            /* a block the header calls for but the caller did not give */
            char zero[4096] = {0};
            for (int64_t left = bytes; left > 0 && ok; left -= (int64_t)sizeof zero)
                ok = fwrite(zero, 1, (size_t)(left < (int64_t)sizeof zero ? left : (int64_t)sizeof zero), f) > 0;
        }
        ok = ok && fwrite(&m, 4, 1, f) == 1;
    }
    if (fclose(f) != 0) ok = 0;
    if (!ok) return snap_err(err, errlen, "write failed", path);
    return PSR_OK;
}

/* ------------------------------------------------------------ epochs */
/* distance of two positions as src/psystem.c:507-518 computes it: float
 * differences (wrapped when periodic), squares added in double, sqrt */
static float epoch_dist(const float *a, const float *b, int periodic, double boxsize) {
    double r = 0.0;
    for (int d = 0; d < 3; d++) {
        float dd = fabsf(a[d] - b[d]);
        if (periodic && dd > 0.5 * boxsize) dd = (float)(boxsize - dd);
        r += dd * dd;
    }
    return (float)sqrt(r);
}

typedef struct {
    uint64_t hash;
    int64_t idx;
} epoch_key;

/* hash ascending, index descending: the order the reference's hash chains
 * visit a bucket (the last inserted first, src/psystem.c:49-58) */
static int epoch_key_cmp(const void *pa, const void *pb) {
    const epoch_key *a = pa, *b = pb;
    if (a->hash != b->hash) return a->hash < b->hash ? -1 : 1;
    return (a->idx < b->idx) - (a->idx > b->idx);
}

/* a fraction x with complement rest, pinned to 0 or 1 when either is out
 * of range (psys_set_lambdaH / psys_set_lambdaHe, src/psystem.h:156-169) */
static double clamp_fraction(double x, double rest) {
    if (x >= 1.0 || rest <= 0.0) return 1.0;
    if (x <= 0.0 || rest >= 1.0) return 0.0;
    return x;
}

#define PSR_PF_INVALID 1
#define PSR_PF_NORMAL 0
#define PSR_IDHASH_MASK ((((uint64_t)1) << 28) - 1)

static void state_swap(psr_state *s, uint64_t *id, int64_t i, int64_t j) {
#define SW(arr, T)                 \
    do {                           \
        const T t_ = (arr)[i];     \
        (arr)[i] = (arr)[j];       \
        (arr)[j] = t_;             \
    } while (0)
    for (int d = 0; d < 3; d++) {
        const float t = s->pos[3 * i + d];
        s->pos[3 * i + d] = s->pos[3 * j + d];
        s->pos[3 * j + d] = t;
    }
    SW(s->lambdaH, double);
    SW(s->lambdaHeI, double);
    SW(s->lambdaHeII, double);
    SW(s->yeMET, float);
    SW(s->mass, float);
    SW(s->sml, float);
    SW(s->rho, float);
    SW(s->ie, float);
    SW(s->flag, int8_t);
    SW(s->yGdepHI, float);
    SW(s->yGdepHeI, float);
    SW(s->yGdepHeII, float);
    SW(s->yGrecHII, float);
    SW(s->yGrecHeII, float);
    SW(s->yGrecHeIII, float);
    SW(s->heat, float);
    SW(s->lasthit, int64_t);
    SW(s->hits, int32_t);
    SW(id, uint64_t);
#undef SW
}

int psr_epoch_load(psr_state *s, uint64_t *id, int64_t npar_max, const psr_epoch_snap *n, const psr_units *u,
                   int32_t h_only, char *err, size_t errlen) {
    if (!s || !id || !n || !u || n->npar < 0 || !n->ye || !n->xHI)
        return PSR_ERR_ARG;
    if (n->npar > npar_max) {
        if (err && errlen) snprintf(err, errlen, "%lld particles beyond npar_max %lld", (long long)n->npar,
                                    (long long)npar_max);
        return PSR_ERR_CAPACITY;
    }
    for (int64_t i = 0; i < n->npar; i++) {
        for (int d = 0; d < 3; d++) s->pos[3 * i + d] = n->pos[3 * i + d];
        s->mass[i] = n->mass[i];
        s->sml[i] = n->sml[i];
        s->rho[i] = n->rho[i];
        s->ie[i] = n->ie[i];
        id[i] = n->id[i];
        const double xHII = 1.0 - n->xHI[i];
        s->lambdaH[i] = clamp_fraction(n->xHI[i], xHII);
        /* electrons beyond hydrogen's go to helium first (src/psystem.c:349-358) */
        double yeMET = fdim(n->ye[i], xHII);
        if (!h_only) {
            double xHeII = yeMET * u->C_HMF / (1 - u->C_HMF) * 4;
            double xHeIII = fdim(xHeII, 1.0);
            xHeII = fmin(1.0, xHeII);
            xHeIII = fmin(1.0 - xHeII, xHeIII);
            const double xHeI = 1.0 - xHeII - xHeIII;
            s->lambdaHeI[i] = clamp_fraction(xHeI, xHeII + xHeIII);
            s->lambdaHeII[i] = clamp_fraction(xHeII, xHeI + xHeIII);
            yeMET = fdim(yeMET, (xHeII + 2 * xHeIII) * (1 - u->C_HMF) / u->C_HMF * 0.25);
        } else {
            s->lambdaHeI[i] = s->lambdaHeII[i] = 0.0;
        }
        s->yeMET[i] = (float)yeMET;
        s->yGdepHI[i] = s->yGdepHeI[i] = s->yGdepHeII[i] = 0.0f;
        s->yGrecHII[i] = s->yGrecHeII[i] = s->yGrecHeIII[i] = 0.0f;
        s->heat[i] = 0.0f;
        s->lasthit[i] = 0;
        s->hits[i] = 0;
        s->flag[i] = PSR_PF_NORMAL;
    }
    s->npar = n->npar;
    return PSR_OK;
}

int psr_epoch_switch(psr_state *s, uint64_t *id, int64_t npar_max, const psr_epoch_snap *n, const psr_units *u,
                     int32_t h_only, int32_t periodic, double boxsize, double prev_duration, int64_t prev_nticks,
                     double tick_time, psr_epoch_stats *out, char *err, size_t errlen) {
    if (!s || !id || !n || !u || s->npar < 0 || s->npar > npar_max || n->npar < 0 || prev_nticks <= 0 ||
        !(tick_time > 0))
        return PSR_ERR_ARG;
    const int64_t nold = s->npar;
    /* the current particles by ID bucket (src/psystem.c:381-382) */
    epoch_key *keys = malloc(sizeof(epoch_key) * (size_t)(nold > 0 ? nold : 1));
    for (int64_t i = 0; i < nold; i++) {
        keys[i].hash = id[i] & PSR_IDHASH_MASK;
        keys[i].idx = i;
    }
    qsort(keys, (size_t)nold, sizeof(epoch_key), epoch_key_cmp);
    for (int64_t i = 0; i < nold; i++) s->flag[i] |= PSR_PF_INVALID;
    int64_t npar = nold, added = 0, lost = 0;
    double distsum = 0.0;
    for (int64_t ip = 0; ip < n->npar; ip++) {
        const uint64_t hash = n->id[ip] & PSR_IDHASH_MASK;
        /* the bucket: keys with this hash */
        int64_t lo = 0, hi = nold;
        while (lo < hi) {
            const int64_t mid = (lo + hi) / 2;
            if (keys[mid].hash < hash) lo = mid + 1;
            else hi = mid;
        }
        int64_t best = -1;
        float mindist = 500.0f;
        for (int64_t k = lo; k < nold && keys[k].hash == hash; k++) {
            const float dst = epoch_dist(&n->pos[3 * ip], &s->pos[3 * keys[k].idx], periodic, boxsize);
            if (dst < mindist) {
                mindist = dst;
                best = keys[k].idx;
            }
        }
        int insert = 0;
        if (best == -1) {
            if (npar == npar_max) {
                lost++;
                continue;
            }
            best = npar++;
            insert = 1;
            added++;
        } else {
            distsum += mindist;
        }
        for (int d = 0; d < 3; d++) s->pos[3 * best + d] = n->pos[3 * ip + d];
        s->mass[best] = n->mass[ip];
        s->sml[best] = n->sml[ip];
        s->rho[best] = n->rho[ip];
        s->ie[best] = n->ie[ip];
        s->flag[best] &= ~PSR_PF_INVALID;
        if (insert) {
            /* a new particle takes the snapshot's ionization state
             * (src/psystem.c:455-485) */
            id[best] = n->id[ip];
            const double xHI = n->xHI[ip], xHII = 1.0 - n->xHI[ip];
            s->lambdaH[best] = clamp_fraction(xHI, xHII);
            double yeMET = fdim(s->yeMET[best], xHII);
            if (!h_only) {
                double xHeII = yeMET * u->C_HMF / (1 - u->C_HMF) * 4;
                double xHeIII = fdim(xHeII, 1.0);
                xHeIII = fmin(1.0, xHeII);
                xHeII = fmin(1.0, xHeII);
                const double xHeI = 1.0 - xHeII - xHeIII;
                s->lambdaHeI[best] = clamp_fraction(xHeI, xHeII + xHeIII);
                s->lambdaHeII[best] = clamp_fraction(xHeII, xHeI + xHeIII);
                yeMET = fdim(yeMET, (xHeII + xHeIII) * (1 - u->C_HMF) / u->C_HMF * 0.25);
            }
            s->yeMET[best] = (float)yeMET;
            s->yGdepHI[best] = s->yGdepHeI[best] = s->yGdepHeII[best] = 0.0f;
            s->yGrecHII[best] = s->yGrecHeII[best] = s->yGrecHeIII[best] = 0.0f;
            /* the reference clears heat and hits of particle 3, its loop
             * counter, not of the new one (src/psystem.c:483-484) */
            if (npar_max > 3) {
                s->heat[3] = 0.0f;
                s->hits[3] = 0;
            }
            s->lasthit[best] = 0;
            s->flag[best] = PSR_PF_NORMAL;
        }
    }
    free(keys);
    /* lasthit in ticks of the new epoch (src/psystem.c:266-272) */
    const double last_ticktime = prev_duration / (double)prev_nticks;
    for (int64_t i = 0; i < npar; i++)
        s->lasthit[i] = (int64_t)(((double)s->lasthit[i] * last_ticktime - prev_duration) / tick_time);
    /* compress (src/psystem.c:109-150): the last valid particles fill the
     * holes from the front */
    int64_t invalid = 0;
    for (int64_t i = 0; i < npar; i++) invalid += (s->flag[i] & PSR_PF_INVALID) != 0;
    int64_t j = npar - 1;
    while (j > 0 && (s->flag[j] & PSR_PF_INVALID)) j--;
    for (int64_t i = 0; i < j; i++) {
        if (s->flag[i] & PSR_PF_INVALID) {
            state_swap(s, id, i, j);
            while ((s->flag[j] & PSR_PF_INVALID) && j > i) j--;
        }
    }
    if (j == 0 && npar > 0 && (s->flag[0] & PSR_PF_INVALID)) j = -1;  /* every particle invalid */
    s->npar = npar > 0 ? j + 1 : 0;
    if (out) {
        out->matched = n->npar - added - lost;
        out->added = added;
        out->lost = lost;
        out->removed = invalid;
        out->mean_shift = n->npar > 0 ? distsum / (double)n->npar : 0.0;
    }
    if (lost) {
        if (err && errlen) snprintf(err, errlen, "%lld unmatched particles beyond npar_max", (long long)lost);
        return PSR_ERR_CAPACITY;
    }
    return PSR_OK;
}
