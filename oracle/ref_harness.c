/*
 * ref_harness.c -- drives the reference's own LSODA (support/lsoda, compiled
 * from /root/reference by oracle/Makefile into oracle/_ref/) the way
 * step_evolve_lsoda does (src/solve.c:149-224), with a caller-supplied
 * right-hand side.  TEST INFRASTRUCTURE ONLY: it checks the oracle's
 * chemistry integrator against the reference integrator.
 */
#include <math.h>
#include <stdlib.h>
#include "lsoda.h"

/* returns the final ctx.state (2 = success); x[7] in/out; clamps as
 * src/solve.c:208-213 */
int ref_lsoda_evolve(int (*fn)(double, double *, double *, void *), void *data, double *x,
                     int mxstep0) {
    double rtol[7] = {1e-4, 1e-4, 1e-4, 1e-4, 1e-4, 1e-4, 1e-4};
    double atol[7] = {1e-7, 1e-7, 1e-7, 1e-7, 1e-7, 1e-7, 1.};
    struct lsoda_context_t ctx = {0};
    struct lsoda_opt_t opt = {0};
    ctx.function = (_lsoda_f)fn;
    ctx.neq = 7;
    opt.ixpr = 0;
    opt.mxstep = mxstep0 > 0 ? mxstep0 : 1000;
    opt.rtol = rtol;
    opt.atol = atol;
    opt.itask = 1;
    lsoda_prepare(&ctx, &opt);
    ctx.data = data;
    ctx.state = 1;
    lsoda_reset(&ctx);
    double t = 0;
    do {
        ctx.state = 1;
        lsoda(&ctx, x, &t, 1.);
        if (ctx.error) {
            free(ctx.error);
            ctx.error = NULL;
        }
        if (ctx.state == -1) opt.mxstep = opt.mxstep * 2;
    } while (ctx.state == -1 && t < 1.);
    int state = ctx.state;
    x[0] = fmin(1., fmax(0., x[0]));
    x[1] = fmin(1., fmax(0., x[1]));
    x[2] = fmin(1. - x[1], fmax(0., x[2]));
    x[3] = fmax(0., x[3]);
    x[4] = fmax(0., x[4]);
    x[5] = fmax(0., x[5]);
    lsoda_free(&ctx);
    return state;
}

/* the reference's Robertson demo problem (support/lsoda/main.c:5-10) as a
 * callback, for the published known-answer test */
int ref_robertson(double t, double *y, double *ydot, void *data) {
    (void)t;
    (void)data;
    ydot[0] = 1.0E4 * y[1] * y[2] - .04E0 * y[0];
    ydot[2] = 3.0E7 * y[1] * y[1];
    ydot[1] = -1.0 * (ydot[0] + ydot[2]);
    return 0;
}

/* integrate Robertson to the 12 output times of support/lsoda/main.c:25-59 */
int ref_robertson_run(double *out /* 12 x 3 */) {
    double atol[3] = {1.0E-6, 1.0E-10, 1.0E-6};
    double rtol[3] = {1.0E-4, 1.0E-4, 1.0E-4};
    double y[3] = {1.0, 0.0, 0.0};
    double t = 0.0, tout = 0.4;
    struct lsoda_opt_t opt = {0};
    opt.ixpr = 0;
    opt.rtol = rtol;
    opt.atol = atol;
    opt.itask = 1;
    struct lsoda_context_t ctx = {0};
    ctx.function = ref_robertson;
    ctx.neq = 3;
    ctx.state = 1;
    lsoda_prepare(&ctx, &opt);
    for (int k = 0; k < 12; k++) {
        lsoda(&ctx, y, &t, tout);
        if (ctx.state <= 0) {
            lsoda_free(&ctx);
            return -1;
        }
        out[3 * k] = y[0];
        out[3 * k + 1] = y[1];
        out[3 * k + 2] = y[2];
        tout *= 10.0;
    }
    lsoda_free(&ctx);
    return 0;
}
