// psr_deposit.hip -- photon deposition along the sorted hit lists on gfx950
// (deposit, src/loop.c:388-608).
//
// One 3-wavefront workgroup per ray, pipelined over segments of 128 sorted
// hits held in an LDS ring of three:
//
//   wave 1 (prep)   gathers the particles of segment s + 1 (sml, mass, x_HI
//                   [, x_HeI, x_HeII]), computes the optical depths, em =
//                   expm1(-tau_sum) and N_HI (src/loop.c:427-477) into the
//                   ring.  em = +0 marks a hit with tau_sum = 0, which the
//                   reference skips (src/loop.c:475-477); a live hit has em
//                   <= -0 (sign bit set).  Slots past the list end get no-op
//                   records (em = +0, N = 0, N_HeI = -1).
//   wave 0 (walk)   one lane runs the transmission recurrence over segment s,
//                       TM -= min(-TM * em, N_HI)          (hydrogen only)
//                   evaluated as TM' = max(TM + TM * em, TM - N_HI): correctly
//                   rounded subtraction is monotone in the subtrahend, so this
//                   is the reference's value bit for bit, three dependent
//                   operations per hit.  It stores TM only every 8 hits (a
//                   checkpoint), and stops at the first group of 8 after which
//                   TM < T0 and which holds a live hit; the cut-off "TM < 0 ||
//                   TM / Nph < 1e-10" (src/loop.c:585) is exactly "TM < T0".
//   wave 2 (apply)  for segment s - 1, each lane recomputes the TM its hit saw
//                   from the checkpoint (at most 7 steps, the same operations,
//                   so the same bits), resolves the exact cut hit in the
//                   stopping group, and computes the ionization / heating
//                   deposits (src/loop.c:486-573): either added to the
//                   particles with f32 atomics, or written as records into
//                   bins by particle range that k_bin_reduce sums in LDS.
//
// The walk is the critical path (about 35 cycles per hit at 2.4 GHz for
// hydrogen); prep and apply run beside it.  The per-hit values never leave
// LDS, and the apply wave never waits on global loads (stores and atomics
// share the in-order vmcnt counter with loads on gfx9): the bin range and
// room of its records are fetched one pipeline iteration ahead.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "psr_device.h"
#include "psr_internal.h"

namespace {

// smallest TM with fl(TM / Nph) >= 1e-10: "TM < 0 || TM / Nph < 1e-10"
// (src/loop.c:585) is then exactly "TM < T0" for Nph > 0, because a
// correctly rounded quotient is monotone in its numerator
__device__ double cutoff_threshold(double Nph) {
    double c = 1e-10 * Nph;
    for (int k = 0; k < 64 && c / Nph >= 1e-10; k++) c = nextafter(c, 0.0);
    for (int k = 0; k < 64 && c / Nph < 1e-10; k++) c = nextafter(c, INFINITY);
    return c;
}

__device__ __forceinline__ double readlane_d(double v, int j) {
    const unsigned long long u = (unsigned long long)__double_as_longlong(v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)u, j);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(u >> 32), j);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

__device__ __forceinline__ bool is_live(double em) { return signbit(em); }

constexpr int FW = 3;                // wavefronts: walk, prep, apply
constexpr int FTHREADS = 64 * FW;
constexpr int FSEG = 128;            // hits per segment
// checkpoint interval of the walk (hits per group): the H + He walk keeps
// one group of 64-byte records in registers and its apply wave recomputes up
// to FG - 1 steps per hit, so 4 (measured 256^3: 9.9 -> 9.3 ms per deposit);
// the hydrogen walk's per-group work weighs more against its short steps, so
// 8 (measured 128^3: 4 costs 152 -> 190 us)
template <bool HE>
struct Grp {
    static constexpr int FG = HE ? 4 : 8;
    static constexpr int FNCK = FSEG / FG;
};

// walk record of one hit (16-byte aligned: one LDS b128 read per pair)
template <bool HE>
struct __attribute__((aligned(16))) WalkRec {
    double em, nhi;
};
template <>
struct __attribute__((aligned(16))) WalkRec<true> {
    double em, nhi, fhi, fhei, fheii, nhei, nheii;
    double tm;  // the several-ray walk stores the TM the hit saw here
};

// what apply needs besides the walk record
template <bool HE>
struct HitAux {
    uint32_t slot;
    float mass;
    double xHI;
};
template <>
struct HitAux<true> {
    uint32_t slot;
    float mass;
    double xHI, xHeI, xHeII;
};

// one step of the transmission recurrence, src/loop.c:479-504, 587
template <bool HE>
__device__ __forceinline__ double walk_step(double TM, const WalkRec<HE> &v);
template <>
__device__ __forceinline__ double walk_step<false>(double TM, const WalkRec<false> &v) {
    // TM - min(-TM * em, N_HI), evaluated as max(TM - (-TM * em), TM - N_HI)
    return fmax(TM + TM * v.em, TM - v.nhi);
}
template <>
__device__ __forceinline__ double walk_step<true>(double TM, const WalkRec<true> &v) {
    // TM - (aHI + (aHeI + aHeII)), each a = min(f * (-TM * em), N): fmin
    // (one instruction) equals the reference's compare and replace for the
    // non-negative operands here, up to the sign of a zero, which leaves TM
    // unchanged.  A hit without helium carries f = N = 0 in the helium
    // terms, so its sum is aHI + 0 = aHI exactly and no select is needed.
    const double ae = -TM * v.em;
    const double a0 = fmin(v.fhi * ae, v.nhi);
    const double a1 = fmin(v.fhei * ae, v.nhei);
    const double a2 = fmin(v.fheii * ae, v.nheii);
    return TM - (a0 + (a1 + a2));
}

template <bool HE>
__device__ __forceinline__ void noop_rec(WalkRec<HE> &r, HitAux<HE> &a) {
    r.em = 0.0;
    r.nhi = 0.0;
    if constexpr (HE) {
        r.fhi = r.fhei = r.fheii = r.nhei = r.nheii = 0.0;
        a.xHeI = a.xHeII = 0.0;
    }
    a.slot = 0;
    a.mass = 0.0f;
    a.xHI = 0.0;
}

// per-ray constants
struct RayConst {
    double sigmaHI, sigmaHeI, sigmaHeII, dEHI, dEHeI, dEHeII, hf_HI, hf_HeI, hf_HeII;
    double sHI_HI, sHeI_HI, sHI_HeI, sHeI_HeI, rCH, rCHe;
    bool secion_terms;
};

template <bool HE>
__device__ __forceinline__ RayConst ray_const(double freq, const DevTables &T, const DevPhys &P) {
    RayConst k;
    k.sigmaHI = xs_get(T, 0, freq);
    k.sigmaHeI = HE ? xs_get(T, 1, freq) : 0.0;
    k.sigmaHeII = HE ? xs_get(T, 2, freq) : 0.0;
    k.dEHI = fdim(freq, P.C_HI_ENERGY);
    k.dEHeI = fdim(freq, P.C_HEI_ENERGY);
    k.dEHeII = fdim(freq, P.C_HEII_ENERGY);
    k.hf_HI = P.C_H_PER_MASS * k.dEHI;
    k.hf_HeI = P.C_HE_PER_MASS * k.dEHeI;
    k.hf_HeII = P.C_HE_PER_MASS * k.dEHeII;
    k.secion_terms = P.secondary_ionization && (k.dEHI > 0 || k.dEHeI > 0) && T.si_n1 > 0;
    k.sHI_HI = k.sHeI_HI = k.sHI_HeI = k.sHeI_HeI = k.rCH = k.rCHe = 0.0;
    if constexpr (HE) {
        k.sHI_HI = k.dEHI / P.C_HI_ENERGY;
        k.sHeI_HI = k.dEHeI / P.C_HI_ENERGY;
        k.sHI_HeI = k.dEHI / P.C_HEI_ENERGY;
        k.sHeI_HeI = k.dEHeI / P.C_HEI_ENERGY;
        k.rCH = 1.0 / P.C_H_PER_MASS;
        k.rCHe = 1.0 / P.C_HE_PER_MASS;
    }
    return k;
}

// particle data of one hit, gathered one pipeline stage before its use
template <bool HE>
struct HitGather {
    uint32_t slot;
    float sml, mass;
    double b, xHI, xHeI, xHeII;
};

template <bool HE>
__device__ __forceinline__ void gather_hit(HitGather<HE> &g, const DevState &S) {
    // one or two 16-byte loads: (sml, mass, xHI) and (xHeI, xHeII)
    const auto *w = gmem((const uint32_t *)__builtin_assume_aligned(S.dv, 32)) +
                    (HE ? 8 * (size_t)g.slot : 4 * (size_t)g.slot);
    const uint32_t a0 = w[0], a1 = w[1], a2 = w[2], a3 = w[3];
    g.sml = __uint_as_float(a0);
    g.mass = __uint_as_float(a1);
    g.xHI = __hiloint2double((int)a3, (int)a2);
    if constexpr (HE) {
        const uint32_t b0 = w[4], b1 = w[5], b2 = w[6], b3 = w[7];
        g.xHeI = __hiloint2double((int)b1, (int)b0);
        g.xHeII = __hiloint2double((int)b3, (int)b2);
    }
}

// the packed view from the particle arrays (after an upload or a change of
// the H / H + He layout)
__global__ void k_build_dview(DevState S, int64_t n, int he) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        if (he) {
            DepViewHe v;
            v.sml = fabsf(S.ps[i].w);
            v.mass = S.mass[i];
            v.xHI = S.lH[i];
            v.xHeI = S.lHeI[i];
            v.xHeII = S.lHeII[i];
            ((DepViewHe *)S.dv)[i] = v;
        } else {
            DepViewH v;
            v.sml = fabsf(S.ps[i].w);
            v.mass = S.mass[i];
            v.xHI = S.lH[i];
            ((DepViewH *)S.dv)[i] = v;
        }
    }
}

// src/loop.c:427-477 for one hit
// (sdepth = sph_depth(b / sml), looked up by the caller so that the table
// loads of several hits are in flight together)
template <bool HE>
__device__ __forceinline__ void prep_hit(const HitGather<HE> &g, double sdepth, const DevPhys &P,
                                         const RayConst &k, WalkRec<HE> &r, HitAux<HE> &a) {
    const uint32_t slot = g.slot;
    const double sml_inv = 1.0 / (double)g.sml;
    const float massf = g.mass;
    const double mass = (double)massf;
    const double xHI = g.xHI;
    const double NH = mass * P.C_H_PER_MASS;
    const double NHe = mass * P.C_HE_PER_MASS;
    const double kernel = sdepth * (sml_inv * sml_inv) * P.scaling_fac2_inv;
    const double NHI = xHI * NH;
    const double tauHI = k.sigmaHI * (kernel * NHI);
    double tausum = tauHI;
    double tauHeI = 0, tauHeII = 0, NHeI = 0, NHeII = 0, xHeI = 0, xHeII = 0;
    const bool he = HE && NHe != 0.0;
    if (HE) {
        xHeI = g.xHeI;
        xHeII = g.xHeII;
    }
    if (he) {
        NHeI = xHeI * NHe;
        NHeII = xHeII * NHe;
        tauHeI = k.sigmaHeI * (kernel * NHeI);
        tausum += tauHeI;
        tauHeII = k.sigmaHeII * (kernel * NHeII);
        tausum += tauHeII;
    }
    const bool live = tausum != 0.0;
    // expm1 of a negative argument is <= -0, so the sign bit marks live
    r.em = live ? expm1(-tausum) : 0.0;
    r.nhi = live ? NHI : 0.0;
    if constexpr (HE) {
        r.fhi = live ? tauHI / tausum : 0.0;
        r.fhei = (live && he) ? tauHeI / tausum : 0.0;
        r.fheii = (live && he) ? tauHeII / tausum : 0.0;
        r.nhei = (live && he) ? NHeI : 0.0;
        r.nheii = (live && he) ? NHeII : 0.0;
        a.xHeI = xHeI;
        a.xHeII = xHeII;
    }
    a.slot = slot;
    a.mass = massf;
    a.xHI = xHI;
}

struct ApplySums {
    int nsat;
    double fiHI, fiHeI, fiHeII, siHI, siHeI;
};

// what one live hit adds to its particle: yGdepHI, heat (and yGdepHeI,
// yGdepHeII), each summed over the reference's separate increments in double
// and rounded once; the hit count is one per record
template <bool HE>
struct DepRec {
    uint32_t slot;
    float dHI, dheat;
};
template <>
struct DepRec<true> {
    uint32_t slot;
    float dHI, dHeI, dHeII, dheat;
};

// the parts of src/loop.c:516-552 that depend on the particle state only:
// x = ye NH / (NH + NHe) (capped at 1) and the secondary-ionization table
// values at (dE_HI, log10 x) and (dE_HeI, log10 x)
struct SecVals {
    double x;
    double vH[3], vHe[3];
};

// x = ye NH / (NH + NHe), capped at 1 (0 without secondary ionization)
template <bool HE>
__device__ __forceinline__ double secion_x(double xHI, double xHeI, double xHeII, float massf, const DevPhys &P) {
    if (!P.secondary_ionization) return 0.0;
    const double mass = (double)massf;
    const double NH = mass * P.C_H_PER_MASS;
    const double NHe = mass * P.C_HE_PER_MASS;
    if (!HE) xHeI = xHeII = 0.0;
    const double ye = (1.0 - xHI) + P.C_HEMF / P.C_HMF * 0.25 * (xHeII + 2 * (1.0 - xHeII - xHeI));
    double x = ye * NH / (NH + NHe);
    if (x > 1.0) x = 1.0;
    return x;
}

template <bool HE>
__device__ __forceinline__ SecVals secion_at(double x, const DevTables &T, const DevPhys &P, const RayConst &k) {
    SecVals sv;
    sv.x = x;
#pragma unroll
    for (int t = 0; t < 3; t++) sv.vH[t] = sv.vHe[t] = 0.0;
    if (P.secondary_ionization) {
        if (x > 0.0 && k.secion_terms) {
            const double log10x = log10(x);
            // phi_HI, phi_HeI, E_h at (dE_HI, log10 x) and at (dE_HeI, log10 x);
            // hydrogen only, deltaHeI is 0 and so is its term for any table value
            tab2_get3(T, k.dEHI, log10x, sv.vH);
            if (HE) tab2_get3(T, k.dEHeI, log10x, sv.vHe);
        }
    }
    return sv;
}

// secion_at from log10 x (-inf where x <= 0); sv.x is then only the gate
// apply_core tests (1 where x > 0)
template <bool HE>
__device__ __forceinline__ SecVals secion_at_lx(double lx, const DevTables &T, const DevPhys &P,
                                                const RayConst &k) {
    SecVals sv;
    sv.x = lx > -INFINITY ? 1.0 : 0.0;
#pragma unroll
    for (int t = 0; t < 3; t++) sv.vH[t] = sv.vHe[t] = 0.0;
    if (P.secondary_ionization && lx > -INFINITY && k.secion_terms) {
        tab2_get3(T, k.dEHI, lx, sv.vH);
        if (HE) tab2_get3(T, k.dEHeI, lx, sv.vHe);
    }
    return sv;
}

template <bool HE>
__device__ __forceinline__ SecVals hit_secion(double xHI, double xHeI, double xHeII, float massf,
                                              const DevTables &T, const DevPhys &P, const RayConst &k) {
    return secion_at<HE>(secion_x<HE>(xHI, xHeI, xHeII, massf, P), T, P, k);
}

// src/loop.c:479-573 for one live hit that saw TM = tm, with the state-only
// values sv of hit_secion
template <bool HE>
__device__ __forceinline__ DepRec<HE> apply_core(double tm, const WalkRec<HE> &v, uint32_t slot, float massf,
                                                 const SecVals &sv, const DevPhys &P, const RayConst &k,
                                                 ApplySums &acc) {
    double yHI = 0.0, yHeI = 0.0, yHeII = 0.0, ht = 0.0;
    const double em = v.em;
    const double NHI = v.nhi;
    const double mass = (double)massf;
    const double NH = mass * P.C_H_PER_MASS;
    const double NHe = mass * P.C_HE_PER_MASS;
    // with helium, the per-hit divisions (six per hit: 1/NH, 1/NHe and the
    // four (delta dE) / E_ion of the secondary terms) become products with
    // per-ray reciprocals and one reciprocal of the mass; the values move
    // by an FP64 rounding, below the f32 the deposits are summed in (the
    // hydrogen-only kernel keeps the divisions, which cost it nothing)
    double NH_inv, NHe_inv;
    if constexpr (HE) {
        const double rmass = 1.0 / mass;
        NH_inv = rmass * k.rCH;
        NHe_inv = rmass * k.rCHe;
    } else {
        NH_inv = 1.0 / NH;
        NHe_inv = 1.0 / NHe;
    }
    // the walk's operations on the TM this hit saw
    const double absorb_est = -tm * em;
    double fHI = 1.0, fHeI = 0.0, fHeII = 0.0, NHeI = 0.0, NHeII = 0.0;
    bool he = false;
    if constexpr (HE) {
        fHI = v.fhi;
        NHeI = v.nhei;
        he = NHe != 0.0;  // src/loop.c:488
        if (he) {
            fHeI = v.fhei;
            fHeII = v.fheii;
            NHeII = v.nheii;
        }
    }
    double aHI = HE ? fHI * absorb_est : absorb_est;
    if (aHI > NHI) {
        aHI = NHI;
        acc.nsat++;
    }
    double aHeI = 0, aHeII = 0;
    if (he) {
        aHeI = fHeI * absorb_est;
        if (aHeI > NHeI) {
            aHeI = NHeI;
            acc.nsat++;
        }
        aHeII = fHeII * absorb_est;
        if (aHeII > NHeII) {
            aHeII = NHeII;
            acc.nsat++;
        }
    }
    const double deltaHI = aHI * NH_inv;
    const double deltaHeI = he ? aHeI * NHe_inv : 0.0;
    const double deltaHeII = he ? aHeII * NHe_inv : 0.0;
    if (P.secondary_ionization && sv.x > 0.0) {
        yHI += deltaHI;
        acc.fiHI += deltaHI * NH;
        if (k.secion_terms) {
            const double sec = HE ? deltaHI * k.sHI_HI * sv.vH[0] + deltaHeI * k.sHeI_HI * sv.vHe[0]
                                  : deltaHI * k.dEHI / P.C_HI_ENERGY * sv.vH[0] +
                                        deltaHeI * k.dEHeI / P.C_HI_ENERGY * sv.vHe[0];
            yHI += sec;
            acc.siHI += sec * NH;
            ht += k.hf_HI * deltaHI * sv.vH[2];
            if (he) {
                const double sec2 = deltaHI * k.sHI_HeI * sv.vH[1] + deltaHeI * k.sHeI_HeI * sv.vHe[1];
                yHeI += deltaHeI + sec2;
                acc.siHeI += sec2 * NHe;
                acc.fiHeI += deltaHeI * NHe;
                ht += k.hf_HeI * deltaHeI * sv.vHe[2];
                yHeII += deltaHeII;
                acc.fiHeII += deltaHeII * NHe;
            }
        } else if (he) {
            yHeI += deltaHeI;
            acc.fiHeI += deltaHeI * NHe;
            yHeII += deltaHeII;
            acc.fiHeII += deltaHeII * NHe;
        }
    } else {
        yHI += deltaHI;
        ht += k.hf_HI * deltaHI;
        if (he) {
            yHeI += deltaHeI;
            yHeII += deltaHeII;
            ht += deltaHeI * k.hf_HeI;
            ht += deltaHeII * k.hf_HeII;
        }
    }
    DepRec<HE> out;
    out.slot = slot;
    out.dHI = (float)yHI;
    out.dheat = (float)ht;
    if constexpr (HE) {
        out.dHeI = (float)yHeI;
        out.dHeII = (float)yHeII;
    }
    return out;
}

// src/loop.c:479-573 for one live hit that saw TM = tm
template <bool HE>
__device__ __forceinline__ DepRec<HE> apply_values(double tm, const WalkRec<HE> &v, const HitAux<HE> &au,
                                                   const DevTables &T, const DevPhys &P, const RayConst &k,
                                                   ApplySums &acc) {
    double xHeI = 0.0, xHeII = 0.0;
    if constexpr (HE) {
        xHeI = au.xHeI;
        xHeII = au.xHeII;
    }
    const SecVals sv = hit_secion<HE>(au.xHI, xHeI, xHeII, au.mass, T, P, k);
    return apply_core<HE>(tm, v, au.slot, au.mass, sv, P, k, acc);
}

// a record added straight to its particle (records beyond a bin's room)
template <bool HE>
__device__ __forceinline__ void rec_atomic(const DevState &S, const DepRec<HE> &r) {
    gadd(&S.yGdepHI[r.slot], r.dHI);
    if (r.dheat != 0.0f) gadd(&S.heat[r.slot], r.dheat);
    if constexpr (HE) {
        if (r.dHeI != 0.0f) gadd(&S.yGdepHeI[r.slot], r.dHeI);
        if (r.dHeII != 0.0f) gadd(&S.yGdepHeII[r.slot], r.dHeII);
    }
    gadd(&S.hits[r.slot], 1);
}

// Deposit records are binned by particle range (BIN_SHIFT: 2048 particles in
// tree order per bin) instead of being added to the particles with scattered
// atomics, which run at about 23 G/s on this part whatever their scope: the
// apply wave reserves room with one returning atomic per (bin, instruction),
// k_bin_reduce then sums each bin in LDS and adds it to the particles.
constexpr int BIN_SHIFT = 11;
constexpr int BIN_P = 1 << BIN_SHIFT;

struct DevBins {
    void *rec;           // records, bin b at [off[b], off[b + 1])
    const uint32_t *off;  // nbins + 1
    uint32_t *fill;       // records claimed this tick (may exceed the room)
};

template <bool HE>
struct DepLds {
    WalkRec<HE> rec[3][FSEG];
    HitAux<HE> aux[3][FSEG];
    double ck[3][Grp<HE>::FNCK];      // TM before hit FG * g of the segment
    uint8_t glive[3][Grp<HE>::FNCK];  // group g of the segment has a live hit
    int32_t stop_seg, stop_grp, last;
    double tm_end;
};

// records of one segment in the apply wave, with their room reserved; the
// reservation's return values are used one pipeline iteration later
template <bool HE>
struct RecCarry {
    static constexpr int HPL = FSEG / 64;
    DepRec<HE> r[HPL];
    bool has[HPL];
    int leader[HPL], rank[HPL];
    uint32_t old[HPL], o0[HPL], o1[HPL];
};

// group this instruction's records by bin; one returning atomic per bin
// (lane `leader`) claims the room, `rank` is the record's place in the group;
// the bin's range is loaded alongside
template <bool HE>
__device__ __forceinline__ void rec_reserve(RecCarry<HE> &c, int h, bool has, const DepRec<HE> &r,
                                            const DevBins &B, int lane) {
    const unsigned long long lt = (1ull << lane) - 1;
    const int b = (int)(r.slot >> BIN_SHIFT);
    unsigned long long rem = __ballot(has);
    unsigned long long gm = 0;
    int leader = lane;
    while (rem) {
        const int l = __ffsll((long long)rem) - 1;
        const int b0 = __builtin_amdgcn_readlane(b, l);
        const unsigned long long m = __ballot(has && b == b0);
        if (has && b == b0) {
            gm = m;
            leader = l;
        }
        rem &= ~m;
    }
    uint32_t old = 0, o0 = 0, o1 = 0;
    if (has && lane == leader) old = atomicAdd(&B.fill[b], (uint32_t)__popcll(gm));
    if (has) {
        o0 = B.off[b];
        o1 = B.off[b + 1];
    }
    c.r[h] = r;
    c.has[h] = has;
    c.leader[h] = leader;
    c.rank[h] = __popcll(gm & lt);
    c.old[h] = old;
    c.o0[h] = o0;
    c.o1[h] = o1;
}

// write the carried records into their bins (or straight to the particles
// when the bin is full)
template <bool HE>
__device__ __forceinline__ void rec_commit(RecCarry<HE> &c, const DevBins &B, const DevState &S) {
#pragma unroll
    for (int h = 0; h < RecCarry<HE>::HPL; h++) {
        const uint32_t base = (uint32_t)__shfl((int)c.old[h], c.leader[h], 64);
        if (c.has[h]) {
            const uint32_t pos = base + (uint32_t)c.rank[h];
            if (pos < c.o1[h] - c.o0[h]) ((DepRec<HE> *)B.rec)[c.o0[h] + pos] = c.r[h];
            else rec_atomic<HE>(S, c.r[h]);
        }
        c.has[h] = false;
    }
}

// BIN: records into the particle-range bins (else f32 atomics into the
// particles); the atomic variant carries no reservation state.  Register
// caps trade a few spills (in the apply wave, off the walk's critical path)
// for resident workgroups: hydrogen only 128 VGPRs, 5 workgroups per CU
// instead of 4 (10% faster on 128^3); H + He 168 VGPRs, 4 instead of 2 (the
// LDS ring's limit; 1.6x faster on 256^3).
template <bool HE, bool BIN>
__global__ __launch_bounds__(FTHREADS, HE ? 3 : 4) void k_deposit(const DevTick *__restrict__ tk, int64_t nray, DevHits H,
                                                      const DevCtx *__restrict__ C, DevStats *__restrict__ st0,
                                                      double *__restrict__ ray_len, double *__restrict__ ray_tm,
                                                      DevBins Bn, const int32_t *halt) {
    __shared__ DepLds<HE> L;
    constexpr int FG = Grp<HE>::FG;
    constexpr int FNCK = Grp<HE>::FNCK;
    if (halted(halt)) return;
    const psr_ray *__restrict__ rays = tk->rays;
    const DevState &S = C->S;
    const DevTables &T = C->T;
    const DevPhys &P = C->P;
    const int lane = threadIdx.x & 63;
    // 0 walk, 1 prep, 2 apply; wave-uniform, so each role runs its own loop
    // (the three loops pass the same number of barriers)
    const int wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    ApplySums acc{0, 0, 0, 0, 0, 0};
    unsigned long long cnt = 0, walked = 0;
    double lost = 0.0;
    constexpr int HPL = FSEG / 64;  // hits per lane and segment
    RecCarry<HE> carry;             // apply wave: records awaiting their room
    if constexpr (BIN) {
#pragma unroll
        for (int h = 0; h < HPL; h++) {
            carry.has[h] = false;
            carry.leader[h] = lane;
            carry.old[h] = 0;
        }
    }
    for (int64_t r = blockIdx.x; r < nray; r += gridDim.x) {
        const int n = H.count[r];
        const double Nph = rays[r].Nph;
        if (n == 0 || Nph == 0) {
            if (threadIdx.x == 0) {
                ray_len[r] = rays[r].length;
                ray_tm[r] = Nph;
            }
            continue;
        }
        const int64_t base = r * (int64_t)H.cap;
        const double T0 = cutoff_threshold(Nph);
        const RayConst kc = ray_const<HE>(rays[r].freq, T, P);
        double TM = (P.no_photon && rays[r].type < 0) ? 0.0 : Nph;  // carried by the walk lane
        const int nseg = (n + FSEG - 1) / FSEG;
        if (threadIdx.x == 0) {
            L.stop_seg = -1;
            L.stop_grp = -1;
            L.last = n - 1;
            L.tm_end = TM;
        }
        __syncthreads();
        if (wv == 1) {
            // three-stage pipeline: sorted slot / b of segment i + 2 and the
            // particle gathers of segment i + 1 are in flight while segment i
            // is computed; each carried set is used, then reloaded (indexes
            // clamped into the list, so every load is in bounds)
            uint32_t a_slot[HPL];
            double a_b[HPL];
            HitGather<HE> gc[HPL];
#pragma unroll
            for (int h = 0; h < HPL; h++) {
                const int k0 = min(h * 64 + lane, n - 1);
                gc[h].slot = H.sslot[base + k0];
                gc[h].b = H.sb[base + k0];
                gather_hit<HE>(gc[h], S);
                const int k1 = min(FSEG + h * 64 + lane, n - 1);
                a_slot[h] = H.sslot[base + k1];
                a_b[h] = H.sb[base + k1];
            }
            // every iteration runs the same straight-line code (also past the
            // list end or the cut, on clamped indexes, into the ring slot no
            // other wave reads in that iteration), so the wait counts stay exact
            for (int i = 0;; i++) {
                const int stop = L.stop_seg;
                if (stop >= 0 ? i > stop + 2 : i > nseg + 1) break;
                const int ring = i % 3;
                double sd[HPL];
#pragma unroll
                for (int h = 0; h < HPL; h++) sd[h] = sph_depth(T, gc[h].b * (1.0 / (double)gc[h].sml));
#pragma unroll
                for (int h = 0; h < HPL; h++) {
                    const int jj = h * 64 + lane;
                    WalkRec<HE> rr;
                    HitAux<HE> au;
                    prep_hit<HE>(gc[h], sd[h], P, kc, rr, au);
                    if (i * FSEG + jj >= n) noop_rec<HE>(rr, au);
                    L.rec[ring][jj] = rr;
                    L.aux[ring][jj] = au;
                    // the walk's test for a cut looks at groups with a live hit
                    static_assert(64 % FG == 0 && 64 / FG <= 64, "groups within a wavefront");
                    const unsigned long long lv = __ballot(is_live(rr.em));
                    if (lane < 64 / FG)
                        L.glive[ring][h * (64 / FG) + lane] = ((lv >> (lane * FG)) & ((1ull << FG) - 1)) != 0;
                }
#pragma unroll
                for (int h = 0; h < HPL; h++) {
                    gc[h].slot = a_slot[h];
                    gc[h].b = a_b[h];
                    gather_hit<HE>(gc[h], S);
                    const int k2 = min((i + 2) * FSEG + h * 64 + lane, n - 1);
                    a_slot[h] = gmem(H.sslot)[base + k2];
                    a_b[h] = gmem(H.sb)[base + k2];
                }
                __syncthreads();
            }
        } else if (wv == 0) {
            // the walk is the critical path: first claim on issue slots
            __builtin_amdgcn_s_setprio(3);
            for (int i = 0;; i++) {
                const int stop = L.stop_seg;
                if (stop >= 0 ? i > stop + 2 : i > nseg + 1) break;
                const int s = i - 1;
                if (s >= 0 && s < nseg && stop < 0 && lane == 0) {
                    const int ring = s % 3;
                    const int ngrp = min(FNCK, (n - s * FSEG + FG - 1) / FG);
                    double t = TM;
                    // double-buffered groups: the next group's records are
                    // read (pinned ahead by a scheduling barrier) while this
                    // group's steps run, so LDS latency stays off the chain.
                    // TM changes only at live hits, so below T0 after a group
                    // with a live hit means the cut is in that group.
                    if constexpr (HE) {
                        // 64-byte records: one group in registers, each
                        // record reloaded with the next group's as soon as
                        // its step has read it
                        WalkRec<HE> A[FG];
#pragma unroll
                        for (int q = 0; q < FG; q++) A[q] = L.rec[ring][q];
                        for (int g = 0; g < ngrp; g++) {
                            const int g1 = min(g + 1, FNCK - 1);
                            L.ck[ring][g] = t;
                            const bool seen = L.glive[ring][g];
#pragma unroll
                            for (int q = 0; q < FG; q++) {
                                t = walk_step<HE>(t, A[q]);
                                A[q] = L.rec[ring][g1 * FG + q];
                            }
                            if (t < T0 && seen) {
                                L.stop_seg = s;
                                L.stop_grp = g;
                                break;
                            }
                        }
                    } else {
                        WalkRec<HE> A[FG], B[FG];
#pragma unroll
                        for (int q = 0; q < FG; q++) A[q] = L.rec[ring][q];
                        for (int g = 0; g < ngrp; g += 2) {
                            const int g1 = min(g + 1, FNCK - 1);
#pragma unroll
                            for (int q = 0; q < FG; q++) B[q] = L.rec[ring][g1 * FG + q];
                            __builtin_amdgcn_sched_barrier(0);
                            L.ck[ring][g] = t;
                            bool seen = L.glive[ring][g];
#pragma unroll
                            for (int q = 0; q < FG; q++) t = walk_step<HE>(t, A[q]);
                            if (t < T0 && seen) {
                                L.stop_seg = s;
                                L.stop_grp = g;
                                break;
                            }
                            if (g + 1 >= ngrp) break;
                            const int g2 = min(g + 2, FNCK - 1);
#pragma unroll
                            for (int q = 0; q < FG; q++) A[q] = L.rec[ring][g2 * FG + q];
                            __builtin_amdgcn_sched_barrier(0);
                            L.ck[ring][g + 1] = t;
                            seen = L.glive[ring][g + 1];
#pragma unroll
                            for (int q = 0; q < FG; q++) t = walk_step<HE>(t, B[q]);
                            if (t < T0 && seen) {
                                L.stop_seg = s;
                                L.stop_grp = g + 1;
                                break;
                            }
                        }
                    }
                    TM = t;
                    if (s == nseg - 1) L.tm_end = t;
                }
                __syncthreads();
            }
            __builtin_amdgcn_s_setprio(0);
        } else {
            for (int i = 0;; i++) {
                const int stop = L.stop_seg;
                if (stop >= 0 ? i > stop + 2 : i > nseg + 1) break;
                // last iteration's records: their room has arrived by now
                if constexpr (BIN) rec_commit<HE>(carry, Bn, S);
                const int s = i - 2;
                if (s >= 0 && s < nseg && (stop < 0 || s <= stop)) {
                    const int ring = s % 3;
                    int lim = min(FSEG, n - s * FSEG) - 1;  // last hit of the segment to apply
                    if (s == stop) {
                        // the exact cut: first live hit of the stopping group
                        // after which TM < T0
                        const int g = L.stop_grp;
                        const int j = g * FG + (lane & (FG - 1));
                        double pre = L.ck[ring][g];
                        for (int q = g * FG; q < j; q++) pre = walk_step<HE>(pre, L.rec[ring][q]);
                        const WalkRec<HE> v = L.rec[ring][j];
                        const double post = walk_step<HE>(pre, v);
                        const unsigned long long m = __ballot(lane < FG && is_live(v.em) && post < T0);
                        const int jc = __ffsll((long long)m) - 1;
                        lim = g * FG + jc;
                        const double tm_cut = readlane_d(post, jc);
                        if (lane == 0) {
                            L.last = s * FSEG + lim;
                            L.tm_end = tm_cut;
                        }
                    }
#pragma unroll
                    for (int h = 0; h < HPL; h++) {
                        const int jj = h * 64 + lane;
                        const WalkRec<HE> v = L.rec[ring][jj];
                        const bool has = jj <= lim && is_live(v.em);
                        DepRec<HE> rr{};
                        if (has) {
                            const int g0 = jj & ~(FG - 1);
                            double pre = L.ck[ring][g0 / FG];
                            for (int q = g0; q < jj; q++) pre = walk_step<HE>(pre, L.rec[ring][q]);
                            rr = apply_values<HE>(pre, v, L.aux[ring][jj], T, P, kc, acc);
                        }
                        if constexpr (BIN) rec_reserve<HE>(carry, h, has, rr, Bn, lane);
                        else if (has) rec_atomic<HE>(S, rr);
                    }
                }
                __syncthreads();
            }
        }
        if (threadIdx.x == 0) {
            const int last = L.last;
            const double tm = L.tm_end;
            ray_len[r] = H.d[base + H.perm[base + last]];
            ray_tm[r] = tm;
            cnt += (unsigned long long)n;
            walked += (unsigned long long)(last + 1);
            lost += tm;
        }
        __syncthreads();
    }
    if constexpr (BIN)
        if (wv == 2) rec_commit<HE>(carry, Bn, S);  // the last segment's records
    // counters: thread 0 holds the ray totals, wave 2 the deposit sums
    DevStats *st = stat_slot(st0);
    if (threadIdx.x == 0 && cnt) {
        atomicAdd(&st->total_deposit_count, cnt);
        atomicAdd(&st->deposit_processed, walked);
        atomicAdd(&st->lost_photon_count_sum, lost);
    }
    if (wv == 2) {
        const int ns = wave_sum_i(acc.nsat);
        const double v1 = wave_sum(acc.fiHI), v2 = wave_sum(acc.siHI);
        double v3 = 0, v4 = 0, v5 = 0;
        if (HE) {
            v3 = wave_sum(acc.fiHeI);
            v4 = wave_sum(acc.fiHeII);
            v5 = wave_sum(acc.siHeI);
        }
        if (lane == 0) {
            if (ns) atomicAdd(&st->saturated_deposit_count, (unsigned long long)ns);
            if (v1 != 0) atomicAdd(&st->first_ionization_HI, v1);
            if (v2 != 0) atomicAdd(&st->secondary_ionization_HI, v2);
            if (HE) {
                if (v3 != 0) atomicAdd(&st->first_ionization_HeI, v3);
                if (v4 != 0) atomicAdd(&st->first_ionization_HeII, v4);
                if (v5 != 0) atomicAdd(&st->secondary_ionization_HeI, v5);
            }
        }
    }
}

// one workgroup per bin: the bin's records summed in LDS, then added to its
// particles (no other kernel touches them meanwhile).  Records are loaded
// RB at a time per thread so that their latencies overlap.
constexpr int RED_THREADS = 1024;
constexpr int RB = 4;

template <bool HE>
__global__ __launch_bounds__(RED_THREADS) void k_bin_reduce(DevBins B, int64_t nbins, int64_t npar,
                                                            const DevCtx *__restrict__ C, const int32_t *halt) {
    __shared__ float s_hi[BIN_P], s_heat[BIN_P];
    __shared__ float s_he1[HE ? BIN_P : 1], s_he2[HE ? BIN_P : 1];
    __shared__ uint32_t s_hits[BIN_P];
    if (halted(halt)) return;
    const DevState &S = C->S;
    const DepRec<HE> *rec = (const DepRec<HE> *)B.rec;
    for (int64_t b = blockIdx.x; b < nbins; b += gridDim.x) {
        for (int p = threadIdx.x; p < BIN_P; p += RED_THREADS) {
            s_hi[p] = 0.0f;
            s_heat[p] = 0.0f;
            s_hits[p] = 0;
            if (HE) {
                s_he1[p] = 0.0f;
                s_he2[p] = 0.0f;
            }
        }
        __syncthreads();
        const uint32_t o0 = B.off[b], room = B.off[b + 1] - o0;
        const uint32_t n = min(B.fill[b], room);
        const uint32_t p0 = (uint32_t)(b << BIN_SHIFT);
        for (uint32_t k0 = threadIdx.x; k0 < n; k0 += RB * RED_THREADS) {
            DepRec<HE> r[RB];
#pragma unroll
            for (int u = 0; u < RB; u++) {
                const uint32_t k = min(k0 + u * RED_THREADS, n - 1);
                r[u] = rec[o0 + k];
            }
#pragma unroll
            for (int u = 0; u < RB; u++) {
                if (k0 + u * RED_THREADS >= n) break;
                const uint32_t p = r[u].slot - p0;
                atomicAdd(&s_hi[p], r[u].dHI);
                if (r[u].dheat != 0.0f) atomicAdd(&s_heat[p], r[u].dheat);
                if constexpr (HE) {
                    if (r[u].dHeI != 0.0f) atomicAdd(&s_he1[p], r[u].dHeI);
                    if (r[u].dHeII != 0.0f) atomicAdd(&s_he2[p], r[u].dHeII);
                }
                atomicAdd(&s_hits[p], 1u);
            }
        }
        __syncthreads();
        for (int p = threadIdx.x; p < BIN_P; p += RED_THREADS) {
            const int64_t g = (int64_t)p0 + p;
            const uint32_t nh = s_hits[p];
            if (g < npar && nh) {
                S.yGdepHI[g] += s_hi[p];
                if (s_heat[p] != 0.0f) S.heat[g] += s_heat[p];
                if (HE) {
                    S.yGdepHeI[g] += s_he1[p];
                    S.yGdepHeII[g] += s_he2[p];
                }
                S.hits[g] += (int32_t)nh;
            }
        }
        __syncthreads();
    }
}

// one workgroup: the room of every bin for the next launch, from this
// launch's claims (claims + 25% + 32, scaled down to the buffer when the sum
// does not fit; rec_cap / nbins each when `uniform`), then the claim counters
// cleared
constexpr int PLAN_THREADS = 1024;

__global__ __launch_bounds__(PLAN_THREADS) void k_bin_plan(uint32_t *__restrict__ fill, int64_t nbins,
                                                           uint64_t rec_cap, int uniform,
                                                           uint32_t *__restrict__ off, const int32_t *halt) {
    __shared__ unsigned long long wsum[PLAN_THREADS / 64];
    if (!uniform && halted(halt)) return;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int64_t per = (nbins + PLAN_THREADS - 1) / PLAN_THREADS;
    const int64_t b0 = min(nbins, (int64_t)t * per), b1 = min(nbins, b0 + per);
    auto want = [&](int64_t b) -> unsigned long long {
        if (uniform) return rec_cap / (unsigned long long)nbins;
        const unsigned long long c = fill[b];
        return c + (c >> 2) + 32;
    };
    unsigned long long sum = 0;
    for (int64_t b = b0; b < b1; b++) sum += want(b);
    // inclusive scan over the wave, then over the waves
    unsigned long long x = sum;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const unsigned long long y = __shfl_up(x, d, 64);
        if (lane >= d) x += y;
    }
    if (lane == 63) wsum[w] = x;
    __syncthreads();
    unsigned long long before = 0, total = 0;
#pragma unroll
    for (int k = 0; k < PLAN_THREADS / 64; k++) {
        const unsigned long long v = wsum[k];
        if (k < w) before += v;
        total += v;
    }
    const double scale = total > rec_cap ? (double)rec_cap / (double)total : 1.0;
    // floor(prefix * scale) is monotone and ends inside the buffer
    unsigned long long run = before + x - sum;
    for (int64_t b = b0; b < b1; b++) {
        off[b] = (uint32_t)((double)run * scale);
        run += want(b);
    }
    if (t == PLAN_THREADS - 1) off[nbins] = (uint32_t)fmin((double)rec_cap, (double)total * scale);
    for (int64_t b = b0; b < b1; b++) fill[b] = 0;
}

// ---- several rays per workgroup
// The kernel above gives each ray a workgroup of three single-wavefront
// roles; each role is bound by its own latency chain, and the LDS ring
// keeps four rays per CU.  Here a workgroup carries MR rays with shorter
// segments (MSEG hits per ray, the same 128 hits per iteration): the MR walks
// run in MR lanes of one wavefront (the same instructions, so MR times the
// walk throughput at the same latency), one wavefront preps the MR * MSEG
// hits, and two apply wavefronts take one hit per lane.  The walk stores the
// TM each hit saw, so the apply reads it instead of stepping again from a
// checkpoint.  Values are those of the kernel above bit for bit (same
// prep_hit, walk_step and apply_values); the cut is tested after each live
// hit.  Rays come in batches of similar hit counts (k_ray_order).
// MR is 2, 4 or 8 (the autotuner times them; one ray with 128-hit segments
// measured 1.6x slower than two on 256^3: a single walking lane leaves the
// walk's latency exposed); MSEG = 128 / MR.
constexpr int ORD_THREADS = 1024;
constexpr int MHITS = 128;            // hits per iteration: MR rays x MSEG
constexpr int MW = 4;                 // wavefronts: walk, prep, apply, apply
constexpr int MTHREADS = 64 * MW;
template <int MR>
struct MShape {
    static_assert(MR == 2 || MR == 4 || MR == 8, "two, four or eight rays per workgroup");
    static constexpr int SEG = MHITS / MR;  // hits per ray and segment
    static constexpr int ROW = SEG + 1;     // padded LDS row (see below)
};

// rays by descending hit count, in buckets of ORD_W hits (a batch's walks
// then end within ORD_W hits of one another; lists longer than ORD_NB
// buckets share the first): a counting sort in one workgroup
constexpr int ORD_NB = 4096, ORD_W = 16;
__device__ __forceinline__ int ord_bucket(int32_t n) { return ORD_NB - 1 - min(n / ORD_W, ORD_NB - 1); }

__global__ __launch_bounds__(ORD_THREADS) void k_ray_order(const int32_t *__restrict__ count, int64_t nray,
                                                           uint32_t *__restrict__ order, const int32_t *halt) {
    static_assert(ORD_NB % ORD_THREADS == 0, "whole runs of buckets per thread");
    constexpr int PER = ORD_NB / ORD_THREADS;
    __shared__ uint32_t hist[ORD_NB];
    __shared__ uint32_t wsum[ORD_THREADS / 64];
    if (halted(halt)) return;
    const int tid = threadIdx.x;
    for (int b = tid; b < ORD_NB; b += ORD_THREADS) hist[b] = 0;
    __syncthreads();
    for (int64_t i = tid; i < nray; i += ORD_THREADS) atomicAdd(&hist[ord_bucket(count[i])], 1u);
    __syncthreads();
    // exclusive scan of the counts: thread t owns buckets [PER t, PER t + PER)
    uint32_t own = 0;
#pragma unroll
    for (int k = 0; k < PER; k++) own += hist[PER * tid + k];
    uint32_t incl = own;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t v = __shfl_up(incl, o, 64);
        if ((tid & 63) >= o) incl += v;
    }
    if ((tid & 63) == 63) wsum[tid >> 6] = incl;
    __syncthreads();
    uint32_t run = incl - own;
    for (int w = 0; w < (tid >> 6); w++) run += wsum[w];
#pragma unroll
    for (int k = 0; k < PER; k++) {
        const uint32_t c = hist[PER * tid + k];
        hist[PER * tid + k] = run;
        run += c;
    }
    __syncthreads();
    for (int64_t i = tid; i < nray; i += ORD_THREADS) {
        const uint32_t pos = atomicAdd(&hist[ord_bucket(count[i])], 1u);
        order[pos] = (uint32_t)i;
    }
}

// what the apply wavefronts read besides the walk record.  H + He: log10 x
// of the secondary-ionization lookup (-inf where x <= 0), computed by the
// prep from the fractions, so the record is 16 bytes instead of 32 and the
// apply's loop holds no logarithm (whose polynomial constants it spilled
// and reloaded every iteration at its register budget; 256^3 H + He deposit
// 6.19 -> 5.66 ms).  Hydrogen: x_HI, and the apply forms x (its prep wave,
// which the short hydrogen walk waits on, takes no extra work)
template <bool HE>
struct MAux {
    uint32_t slot;
    float mass;
    double xHI;
};
template <>
struct MAux<true> {
    uint32_t slot;
    float mass;
    double lx;
};

// a ray's row of records (and of TMs) is padded by one entry, so that the
// MR walking lanes, which read the same record index of different rays, hit
// different LDS banks.  The H + He record carries the TM its hit saw (the
// walk writes it); hydrogen's 16-byte record keeps a separate TM row.
template <bool HE, int MR>
struct MLds {
    static constexpr int MROW = MShape<MR>::ROW;
    WalkRec<HE> rec[3][MR * MROW];
    MAux<HE> aux[3][MHITS];
    double tmh[3][HE ? 1 : MR * MROW];  // TM before each hit (walk -> apply), hydrogen
    RayConst kc[MR];           // per-ray constants (cross sections, energies)
    int64_t ray[MR];           // ray index, -1 for an empty place in the batch
    int32_t n[MR];             // hits of the ray (0: nothing to deposit)
    int32_t endseg[MR];        // last segment to deposit (the cut's, once found)
    int32_t last[MR];          // last hit deposited
    double tm_end[MR];
};

// hit k of a ray's list for the prep wavefront: ray place p, hit index in
// the list (clamped into it; a ray without hits reads particle 0)
template <bool HE, int MR>
__device__ __forceinline__ void mr_index(HitGather<HE> &g, const DevHits &H, const MLds<HE, MR> &L, int p, int k) {
    const int n = L.n[p];
    if (n > 0) {
        const int64_t at = L.ray[p] * (int64_t)H.cap + min(k, n - 1);
        g.slot = gmem(H.sslot)[at];
        g.b = gmem(H.sb)[at];
    } else {
        g.slot = 0;
        g.b = 0.0;
    }
}

template <bool HE, int MR>
__global__ __launch_bounds__(MTHREADS, MR == 2 ? 4 : 3) void k_deposit_multi(const DevTick *__restrict__ tk, int64_t nray,
                                                               DevHits H, const DevCtx *__restrict__ C,
                                                               DevStats *__restrict__ st0,
                                                               double *__restrict__ ray_len,
                                                               double *__restrict__ ray_tm, const int32_t *halt) {
    constexpr int MSEG = MShape<MR>::SEG, MROW = MShape<MR>::ROW;
    __shared__ MLds<HE, MR> L;
    if (halted(halt)) return;
    const psr_ray *__restrict__ rays = tk->rays;
    const DevState &S = C->S;
    const DevTables &T = C->T;
    const DevPhys &P = C->P;
    const int lane = threadIdx.x & 63;
    const int wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    ApplySums acc{0, 0, 0, 0, 0, 0};
    unsigned long long cnt = 0, walked = 0;
    double lost = 0.0;
    const int64_t nbatch = (nray + MR - 1) / MR;
    for (int64_t bt = blockIdx.x; bt < nbatch; bt += gridDim.x) {
        if (threadIdx.x < MR) {
            const int p = threadIdx.x;
            const int64_t q = bt * MR + p;
            int64_t r = -1;
            int n = 0;
            if (q < nray) {
                r = H.order[q];
                n = H.count[r];
                if (rays[r].Nph == 0 || n == 0) {
                    // nothing to deposit (src/loop.c:408): the ray keeps its length
                    ray_len[r] = rays[r].length;
                    ray_tm[r] = rays[r].Nph;
                    n = 0;
                }
            }
            L.kc[p] = ray_const<HE>(n > 0 ? rays[r].freq : 0.0, T, P);
            L.ray[p] = r;
            L.n[p] = n;
            L.endseg[p] = n > 0 ? (n - 1) / MSEG : -1;
            L.last[p] = n - 1;
            L.tm_end[p] = 0.0;
        }
        __syncthreads();
        if (wv == 1) {
            // prep: hit j = h * 64 + lane of the iteration is hit (j % MSEG)
            // of ray place j / MSEG.  Four stages, one iteration apart: the
            // slot / b of segment i + 3, the particle gathers of segment
            // i + 2, the SPH table value of segment i + 1 (from the gathered
            // b / sml) are in flight while segment i is computed (indexes
            // clamped into each list)
            constexpr int HPL = 2;
            uint32_t a_slot[HPL];
            double a_b[HPL];
            HitGather<HE> g1[HPL], g2[HPL];   // gathered: segment i (+ sd), segment i + 1
            double sd[HPL];
#pragma unroll
            for (int h = 0; h < HPL; h++) {
                const int j = h * 64 + lane, p = j / MSEG, k = j % MSEG;
                mr_index<HE, MR>(g1[h], H, L, p, k);
                gather_hit<HE>(g1[h], S);
                mr_index<HE, MR>(g2[h], H, L, p, MSEG + k);
                gather_hit<HE>(g2[h], S);
                HitGather<HE> t1;
                mr_index<HE, MR>(t1, H, L, p, 2 * MSEG + k);
                a_slot[h] = t1.slot;
                a_b[h] = t1.b;
                sd[h] = sph_depth(T, g1[h].b * (1.0 / (double)g1[h].sml));
            }
            for (int i = 0;; i++) {
                int mend = -1;
#pragma unroll
                for (int p = 0; p < MR; p++) mend = max(mend, L.endseg[p]);
                if (i > mend + 2) break;
                const int ring = i % 3;
#pragma unroll
                for (int h = 0; h < HPL; h++) {
                    const int j = h * 64 + lane, p = j / MSEG, k = j % MSEG;
                    // a segment past its ray's last (the list's end, or the
                    // walk's cut, which only lowers it) is never walked or
                    // applied: the pipeline's last two iterations and a cut
                    // ray's rest cost no prep
                    if (i > L.endseg[p]) continue;
                    WalkRec<HE> rr;
                    HitAux<HE> au;
                    const RayConst &kc = L.kc[p];
                    prep_hit<HE>(g1[h], sd[h], P, kc, rr, au);
                    if (i * MSEG + k >= L.n[p]) noop_rec<HE>(rr, au);
                    L.rec[ring][p * MROW + k] = rr;
                    MAux<HE> ma;
                    ma.slot = au.slot;
                    ma.mass = au.mass;
                    if constexpr (HE) {
                        const double x = secion_x<HE>(au.xHI, au.xHeI, au.xHeII, au.mass, P);
                        // (-inf: x <= 0; without a table only the gate x > 0
                        // counts, so no logarithm is taken)
                        ma.lx = x > 0.0 ? (T.si_n1 > 0 ? log10(x) : 0.0) : -INFINITY;
                    } else {
                        ma.xHI = au.xHI;
                    }
                    L.aux[ring][j] = ma;
                }
#pragma unroll
                for (int h = 0; h < HPL; h++) {
                    const int j = h * 64 + lane, p = j / MSEG, k = j % MSEG;
                    g1[h] = g2[h];
                    sd[h] = sph_depth(T, g1[h].b * (1.0 / (double)g1[h].sml));
                    g2[h].slot = a_slot[h];
                    g2[h].b = a_b[h];
                    gather_hit<HE>(g2[h], S);
                    HitGather<HE> t2;
                    mr_index<HE, MR>(t2, H, L, p, (i + 3) * MSEG + k);
                    a_slot[h] = t2.slot;
                    a_b[h] = t2.b;
                }
                __syncthreads();
            }
        } else if (wv == 0) {
            // the walks: lane p < MR walks ray place p, one segment behind the
            // prep; the next record is read while a step runs
            __builtin_amdgcn_s_setprio(3);
            const int p = lane < MR ? lane : 0;
            double t = 0.0, T0 = 0.0;
            if (lane < MR && L.n[p] > 0) {
                const int64_t r = L.ray[p];
                const double Nph = rays[r].Nph;
                T0 = cutoff_threshold(Nph);
                t = (P.no_photon && rays[r].type < 0) ? 0.0 : Nph;
            }
            for (int i = 0;; i++) {
                int mend = -1;
#pragma unroll
                for (int q = 0; q < MR; q++) mend = max(mend, L.endseg[q]);
                if (i > mend + 2) break;
                const int s = i - 1;
                if (lane < MR && s >= 0 && s <= L.endseg[p]) {
                    // groups of WG records in registers, each record reloaded
                    // with the next group's as soon as its step has read it;
                    // past the list end the records are no-ops (TM unchanged).
                    // TM changes only at live hits, so the cut (the first
                    // live hit after which TM < T0) is looked for only in a
                    // group that ends below T0, from the TMs after each step
                    constexpr int WG = 4;
                    const int ring = s % 3;
                    const int kn = min(MSEG, L.n[p] - s * MSEG);
                    WalkRec<HE> *R = &L.rec[ring][p * MROW];
                    double *TMH = &L.tmh[ring][HE ? 0 : p * MROW];
                    WalkRec<HE> A[WG];
#pragma unroll
                    for (int q = 0; q < WG; q++) A[q] = R[q];
                    for (int g = 0; g * WG < kn; g++) {
                        double post[WG];
                        bool lv[WG];
                        const int g1 = min(g + 1, MSEG / WG - 1);
#pragma unroll
                        for (int q = 0; q < WG; q++) {
                            if constexpr (HE) R[g * WG + q].tm = t;
                            else TMH[g * WG + q] = t;
                            lv[q] = is_live(A[q].em);
                            t = walk_step<HE>(t, A[q]);
                            post[q] = t;
                            A[q] = R[g1 * WG + q];
                        }
                        if (t < T0) {
                            int qc = -1;
#pragma unroll
                            for (int q = WG - 1; q >= 0; q--)
                                if (lv[q] && post[q] < T0) qc = q;
                            if (qc >= 0) {
                                L.endseg[p] = s;
                                L.last[p] = s * MSEG + g * WG + qc;
#pragma unroll
                                for (int q = 0; q < WG; q++)
                                    if (q == qc) t = post[q];
                                break;
                            }
                        }
                    }
                    L.tm_end[p] = t;
                }
                __syncthreads();
            }
            __builtin_amdgcn_s_setprio(0);
        } else {
            // apply: one hit per lane of segment i - 2 (its walk and cut are
            // done).  The secondary-ionization values depend on the particle
            // state only, so those of segment i - 1 (prepped one iteration
            // ago) are looked up here and used next iteration: their loads
            // are in flight meanwhile, and the wait for them never covers
            // this iteration's atomics (issued after them)
            const int j = (wv - 2) * 64 + lane, p = j / MSEG, k = j % MSEG;
            const RayConst &kc = L.kc[p];
            SecVals nxt{};
            for (int i = 0;; i++) {
                int mend = -1;
#pragma unroll
                for (int q = 0; q < MR; q++) mend = max(mend, L.endseg[q]);
                if (i > mend + 2) break;
                const SecVals cur = nxt;
                const int s1 = i - 1;
                if (s1 >= 0 && L.n[p] > 0) {
                    const MAux<HE> &au = L.aux[s1 % 3][j];
                    if constexpr (HE)
                        nxt = secion_at_lx<HE>(au.lx, T, P, kc);
                    else
                        nxt = secion_at<HE>(secion_x<HE>(au.xHI, 0.0, 0.0, au.mass, P), T, P, kc);
                }
                const int s = i - 2;
                if (s >= 0 && s <= L.endseg[p] && s * MSEG + k <= L.last[p]) {
                    const int ring = s % 3;
                    const WalkRec<HE> v = L.rec[ring][p * MROW + k];
                    if (is_live(v.em)) {
                        const MAux<HE> &au = L.aux[ring][j];
                        double tm;
                        if constexpr (HE) tm = v.tm;
                        else tm = L.tmh[ring][p * MROW + k];
                        const DepRec<HE> rr = apply_core<HE>(tm, v, au.slot, au.mass, cur, P, kc, acc);
                        rec_atomic<HE>(S, rr);
                    }
                }
                __syncthreads();
            }
        }
        if (threadIdx.x < MR && L.n[threadIdx.x] > 0) {
            const int p = threadIdx.x;
            const int64_t r = L.ray[p];
            const int last = L.last[p];
            const int64_t base = r * (int64_t)H.cap;
            ray_len[r] = H.d[base + H.perm[base + last]];
            ray_tm[r] = L.tm_end[p];
            cnt += (unsigned long long)L.n[p];
            walked += (unsigned long long)(last + 1);
            lost += L.tm_end[p];
        }
        __syncthreads();
    }
    // counters: wave 0 holds the ray totals, waves 2 and 3 the deposit sums
    if (wv == 0) {
        const unsigned long long c1 = wave_sum_u64(cnt), c2 = wave_sum_u64(walked);
        const double l1 = wave_sum(lost);
        if (lane == 0 && c1) {
            DevStats *st = stat_slot(st0);
            atomicAdd(&st->total_deposit_count, c1);
            atomicAdd(&st->deposit_processed, c2);
            atomicAdd(&st->lost_photon_count_sum, l1);
        }
    } else if (wv >= 2) {
        const int ns = wave_sum_i(acc.nsat);
        const double v1 = wave_sum(acc.fiHI), v2 = wave_sum(acc.siHI);
        double v3 = 0, v4 = 0, v5 = 0;
        if (HE) {
            v3 = wave_sum(acc.fiHeI);
            v4 = wave_sum(acc.fiHeII);
            v5 = wave_sum(acc.siHeI);
        }
        if (lane == 0) {
            DevStats *st = stat_slot(st0);
            if (ns) atomicAdd(&st->saturated_deposit_count, (unsigned long long)ns);
            if (v1 != 0) atomicAdd(&st->first_ionization_HI, v1);
            if (v2 != 0) atomicAdd(&st->secondary_ionization_HI, v2);
            if (HE) {
                if (v3 != 0) atomicAdd(&st->first_ionization_HeI, v3);
                if (v4 != 0) atomicAdd(&st->first_ionization_HeII, v4);
                if (v5 != 0) atomicAdd(&st->secondary_ionization_HeI, v5);
            }
        }
    }
}

}  // namespace

// bins sized for the context's particles and hit slots; reallocated (after
// the queued ticks finished) when either grows
static int ensure_bins(psr_ctx *c, bool he) {
    const int64_t nbins = (c->npar + BIN_P - 1) / BIN_P;
    const size_t rsz = he ? sizeof(DepRec<true>) : sizeof(DepRec<false>);
    // room for every hit of a tick at the slot capacity, or 16 per particle
    uint64_t rec_cap = std::min<uint64_t>((uint64_t)c->hit_slots * (uint64_t)c->H.cap,
                                          16 * (uint64_t)c->npar + (1u << 20));
    rec_cap = std::max<uint64_t>(rec_cap, 1u << 20);
    rec_cap = std::min<uint64_t>(rec_cap, 0xF0000000u);
    if (nbins <= c->bin_n && rec_cap * rsz <= c->bin_rec_bytes && (int)he <= c->bin_he) return PSR_OK;
    PSR_HIP(c, hipStreamSynchronize(c->stream));
    if (c->d_bin_rec) PSR_HIP(c, hipFree(c->d_bin_rec));
    if (c->d_bin_off) PSR_HIP(c, hipFree(c->d_bin_off));
    c->d_bin_rec = nullptr;
    c->d_bin_off = nullptr;
    c->bin_n = 0;
    PSR_HIP(c, hipMalloc(&c->d_bin_rec, rec_cap * rsz));
    // off[nbins + 1], fill[nbins]
    PSR_HIP(c, hipMalloc(&c->d_bin_off, sizeof(uint32_t) * (2 * (size_t)nbins + 1)));
    c->bin_n = nbins;
    c->bin_rec_bytes = rec_cap * rsz;
    c->bin_rec_cap = rec_cap;
    c->bin_he = he;
    k_bin_plan<<<1, PLAN_THREADS, 0, c->stream>>>(c->d_bin_off + nbins + 1, nbins, rec_cap, 1, c->d_bin_off,
                                                  nullptr);
    PSR_HIP(c, hipGetLastError());
    return PSR_OK;
}

int psr_prepare_deposit(psr_ctx *c, int mode) {
    const int he = !c->F.h_only;
    if (c->have_state && c->npar > 0 && (!c->dview_valid || c->dview_he != he)) {
        c->dview_he = he;
        c->ctx_dirty = true;
        k_build_dview<<<(unsigned)std::min<int64_t>((c->npar + 255) / 256, 16384), 256, 0, c->stream>>>(
            c->state(), c->npar, he);
        PSR_HIP(c, hipGetLastError());
        c->dview_valid = true;
    }
    int rc = psr_sync_dev_ctx(c);
    if (rc) return rc;
    return mode == PSR_DEP_WG_BINNED ? ensure_bins(c, !c->F.h_only) : PSR_OK;
}

// the several-ray deposits' batch order (rays by descending hit count); it
// reads only the tick's hit counts, so the tick launches it right after the
// sort, before waiting for the previous tick's update
int psr_launch_ray_order(psr_ctx *c, int64_t nray) {
    k_ray_order<<<1, ORD_THREADS, 0, c->stream>>>(c->H.count, nray, c->H.order, &c->d_stats->flags);
    PSR_HIP(c, hipGetLastError());
    return PSR_OK;
}

// the rays of the tick come from the tick record (ray_first is kept for the
// call's shape checks); the several-ray modes need psr_launch_ray_order first
int psr_launch_deposit(psr_ctx *c, int64_t, int64_t nray, int mode) {
    int rc = psr_prepare_deposit(c, mode);
    if (rc) return rc;
    const bool he = !c->F.h_only;
    const DevTick *rays = c->d_tick;
    if (mode >= PSR_DEP_MULTI) {
        const int mr = psr_dep_rays(mode);
        const int32_t *halt = &c->d_stats->flags;
        PSR_EV(c, 2, 0);
        const unsigned gm = (unsigned)std::max<int64_t>(1, std::min<int64_t>((nray + mr - 1) / mr, 65535 * 4));
#define DEP_M(HE_, MR_)                                                                                 \
    k_deposit_multi<HE_, MR_><<<gm, MTHREADS, 0, c->stream>>>(rays, nray, c->H, c->d_ctx, c->d_stats, \
                                                              c->d_ray_len, c->d_ray_tm, halt)
        if (he) {
            if (mr == 2) DEP_M(true, 2);
            else if (mr == 4) DEP_M(true, 4);
            else DEP_M(true, 8);
        } else {
            if (mr == 2) DEP_M(false, 2);
            else if (mr == 4) DEP_M(false, 4);
            else DEP_M(false, 8);
        }
#undef DEP_M
        PSR_HIP(c, hipGetLastError());
        PSR_EV(c, 2, 1);
        return PSR_OK;
    }
    const bool binned = mode == PSR_DEP_WG_BINNED;
    const unsigned grid = (unsigned)std::max<int64_t>(1, std::min<int64_t>(nray, 65535 * 4));
    const int32_t *halt = &c->d_stats->flags;
    const int64_t nbins = binned ? c->bin_n : 0;
    uint32_t *off = binned ? c->d_bin_off : nullptr, *fill = binned ? off + nbins + 1 : nullptr;
    const DevBins Bn{binned ? c->d_bin_rec : nullptr, off, fill};
    const unsigned gbin = (unsigned)std::max<int64_t>(1, std::min<int64_t>(nbins, 65535 * 4));
    PSR_EV(c, 2, 0);
#define DEP_K(HE_, BIN_)                                                                                \
    k_deposit<HE_, BIN_><<<grid, FTHREADS, 0, c->stream>>>(rays, nray, c->H, c->d_ctx, c->d_stats, c->d_ray_len, \
                                                           c->d_ray_tm, Bn, halt)
    if (he) {
        if (binned) DEP_K(true, true);
        else DEP_K(true, false);
    } else {
        if (binned) DEP_K(false, true);
        else DEP_K(false, false);
    }
#undef DEP_K
    PSR_HIP(c, hipGetLastError());
    if (binned) {
        if (he) k_bin_reduce<true><<<gbin, RED_THREADS, 0, c->stream>>>(Bn, nbins, c->npar, c->d_ctx, halt);
        else k_bin_reduce<false><<<gbin, RED_THREADS, 0, c->stream>>>(Bn, nbins, c->npar, c->d_ctx, halt);
        PSR_HIP(c, hipGetLastError());
        k_bin_plan<<<1, PLAN_THREADS, 0, c->stream>>>(fill, nbins, c->bin_rec_cap, 0, off, halt);
        PSR_HIP(c, hipGetLastError());
    }
    PSR_EV(c, 2, 1);
    return PSR_OK;
}
