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
