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
//                   stopping group, and scatters the ionization / heating
//                   deposits with f32 atomics (src/loop.c:486-573).
//
// The walk is the critical path (about 35 cycles per hit at 2.4 GHz); prep
// and apply run beside it.  The per-hit values never leave LDS, and the wave
// that issues atomics never waits on global loads (stores and atomics share
// the in-order vmcnt counter with loads on gfx9).
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
constexpr int FG = 8;                // checkpoint interval of the walk
constexpr int FNCK = FSEG / FG;

// walk record of one hit (16-byte aligned: one LDS b128 read per pair)
template <bool HE>
struct __attribute__((aligned(16))) WalkRec {
    double em, nhi;
};
template <>
struct __attribute__((aligned(16))) WalkRec<true> {
    double em, nhi, fhi, fhei, fheii, nhei, nheii, pad;
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
    const double ae = -TM * v.em;
    double a0 = v.fhi * ae;
    a0 = (a0 > v.nhi) ? v.nhi : a0;
    double a1 = v.fhei * ae;
    a1 = (a1 > v.nhei) ? v.nhei : a1;
    double a2 = v.fheii * ae;
    a2 = (a2 > v.nheii) ? v.nheii : a2;
    const double absorb = (v.nhei >= 0.0) ? a0 + (a1 + a2) : a0;
    return TM - absorb;
}

template <bool HE>
__device__ __forceinline__ void noop_rec(WalkRec<HE> &r, HitAux<HE> &a) {
    r.em = 0.0;
    r.nhi = 0.0;
    if constexpr (HE) {
        r.fhi = r.fhei = r.fheii = r.nheii = 0.0;
        r.nhei = -1.0;
        a.xHeI = a.xHeII = 0.0;
    }
    a.slot = 0;
    a.mass = 0.0f;
    a.xHI = 0.0;
}

// per-ray constants
struct RayConst {
    double sigmaHI, sigmaHeI, sigmaHeII, dEHI, dEHeI, dEHeII, hf_HI, hf_HeI, hf_HeII;
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
    g.sml = gmem(S.ps)[g.slot].w;
    g.mass = gmem(S.mass)[g.slot];
    g.xHI = gmem(S.lH)[g.slot];
    if constexpr (HE) {
        g.xHeI = gmem(S.lHeI)[g.slot];
        g.xHeII = gmem(S.lHeII)[g.slot];
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
        r.nhei = (live && he) ? NHeI : -1.0;  // -1: no helium terms
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

// src/loop.c:479-573 for one live hit that saw TM = tm
template <bool HE>
__device__ __forceinline__ void apply_hit(double tm, const WalkRec<HE> &v, const HitAux<HE> &au,
                                          const DevState &S, const DevTables &T, const DevPhys &P,
                                          const RayConst &k, ApplySums &acc) {
    const uint32_t slot = au.slot;
    const double em = v.em;
    const double NHI = v.nhi;
    const double mass = (double)au.mass;
    const double NH = mass * P.C_H_PER_MASS;
    const double NHe = mass * P.C_HE_PER_MASS;
    const double NH_inv = 1.0 / NH;
    const double NHe_inv = 1.0 / NHe;
    // the walk's operations on the TM this hit saw
    const double absorb_est = -tm * em;
    double fHI = 1.0, fHeI = 0.0, fHeII = 0.0, NHeI = 0.0, NHeII = 0.0;
    bool he = false;
    if constexpr (HE) {
        fHI = v.fhi;
        NHeI = v.nhei;
        he = NHeI >= 0.0;
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
    double x = 0.0;
    if (P.secondary_ionization) {
        const double xHI = au.xHI;
        double xHeI = 0.0, xHeII = 0.0;
        if constexpr (HE) {
            xHeI = au.xHeI;
            xHeII = au.xHeII;
        }
        const double ye = (1.0 - xHI) + P.C_HEMF / P.C_HMF * 0.25 * (xHeII + 2 * (1.0 - xHeII - xHeI));
        x = ye * NH / (NH + NHe);
        if (x > 1.0) x = 1.0;
    }
    if (P.secondary_ionization && x > 0.0) {
        atomicAdd(&S.yGdepHI[slot], (float)deltaHI);
        acc.fiHI += deltaHI * NH;
        if (k.secion_terms) {
            const double log10x = log10(x);
            const double sec = deltaHI * k.dEHI / P.C_HI_ENERGY * tab2_get(T, T.si_phi_hi, k.dEHI, log10x) +
                               deltaHeI * k.dEHeI / P.C_HI_ENERGY * tab2_get(T, T.si_phi_hi, k.dEHeI, log10x);
            atomicAdd(&S.yGdepHI[slot], (float)sec);
            acc.siHI += sec * NH;
            atomicAdd(&S.heat[slot], (float)(k.hf_HI * deltaHI * tab2_get(T, T.si_eh, k.dEHI, log10x)));
            if (he) {
                const double sec2 =
                    deltaHI * k.dEHI / P.C_HEI_ENERGY * tab2_get(T, T.si_phi_hei, k.dEHI, log10x) +
                    deltaHeI * k.dEHeI / P.C_HEI_ENERGY * tab2_get(T, T.si_phi_hei, k.dEHeI, log10x);
                atomicAdd(&S.yGdepHeI[slot], (float)(deltaHeI + sec2));
                acc.siHeI += sec2 * NHe;
                acc.fiHeI += deltaHeI * NHe;
                atomicAdd(&S.heat[slot], (float)(k.hf_HeI * deltaHeI * tab2_get(T, T.si_eh, k.dEHeI, log10x)));
                atomicAdd(&S.yGdepHeII[slot], (float)deltaHeII);
                acc.fiHeII += deltaHeII * NHe;
            }
        } else if (he) {
            atomicAdd(&S.yGdepHeI[slot], (float)deltaHeI);
            acc.fiHeI += deltaHeI * NHe;
            atomicAdd(&S.yGdepHeII[slot], (float)deltaHeII);
            acc.fiHeII += deltaHeII * NHe;
        }
    } else {
        atomicAdd(&S.yGdepHI[slot], (float)deltaHI);
        if (k.hf_HI != 0.0) atomicAdd(&S.heat[slot], (float)(k.hf_HI * deltaHI));
        if (he) {
            atomicAdd(&S.yGdepHeI[slot], (float)deltaHeI);
            atomicAdd(&S.yGdepHeII[slot], (float)deltaHeII);
            if (k.hf_HeI != 0.0) atomicAdd(&S.heat[slot], (float)(deltaHeI * k.hf_HeI));
            if (k.hf_HeII != 0.0) atomicAdd(&S.heat[slot], (float)(deltaHeII * k.hf_HeII));
        }
    }
    atomicAdd(&S.hits[slot], 1);
}

template <bool HE>
struct DepLds {
    WalkRec<HE> rec[3][FSEG];
    HitAux<HE> aux[3][FSEG];
    double ck[3][FNCK];       // TM before hit FG * g of the segment
    int32_t stop_seg, stop_grp, last;
    double tm_end;
};

template <bool HE>
__global__ __launch_bounds__(FTHREADS) void k_deposit(const psr_ray *__restrict__ rays, int64_t nray, DevHits H,
                                                      const DevCtx *__restrict__ C, DevStats *__restrict__ st0,
                                                      double *__restrict__ ray_len, double *__restrict__ ray_tm,
                                                      const int32_t *halt) {
    __shared__ DepLds<HE> L;
    if (halted(halt)) return;
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
                    WalkRec<HE> A[FG], B[FG];
#pragma unroll
                    for (int q = 0; q < FG; q++) A[q] = L.rec[ring][q];
                    for (int g = 0; g < ngrp; g += 2) {
                        const int g1 = min(g + 1, FNCK - 1);
#pragma unroll
                        for (int q = 0; q < FG; q++) B[q] = L.rec[ring][g1 * FG + q];
                        __builtin_amdgcn_sched_barrier(0);
                        L.ck[ring][g] = t;
                        bool seen = false;
#pragma unroll
                        for (int q = 0; q < FG; q++) {
                            seen = seen | is_live(A[q].em);
                            t = walk_step<HE>(t, A[q]);
                        }
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
                        seen = false;
#pragma unroll
                        for (int q = 0; q < FG; q++) {
                            seen = seen | is_live(B[q].em);
                            t = walk_step<HE>(t, B[q]);
                        }
                        if (t < T0 && seen) {
                            L.stop_seg = s;
                            L.stop_grp = g + 1;
                            break;
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
                    for (int h = 0; h < FSEG / 64; h++) {
                        const int jj = h * 64 + lane;
                        if (jj <= lim) {
                            const WalkRec<HE> v = L.rec[ring][jj];
                            if (is_live(v.em)) {
                                const int g0 = jj & ~(FG - 1);
                                double pre = L.ck[ring][g0 / FG];
                                for (int q = g0; q < jj; q++) pre = walk_step<HE>(pre, L.rec[ring][q]);
                                apply_hit<HE>(pre, v, L.aux[ring][jj], S, T, P, kc, acc);
                            }
                        }
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

}  // namespace

int psr_launch_deposit(psr_ctx *c, int64_t ray_first, int64_t nray) {
    int rc = psr_sync_dev_ctx(c);
    if (rc) return rc;
    const psr_ray *rays = c->d_rays + ray_first;
    const unsigned grid = (unsigned)std::max<int64_t>(1, std::min<int64_t>(nray, 65535 * 4));
    const int32_t *halt = &c->d_stats->flags;
    PSR_EV(c, 2, 0);
    if (!c->F.h_only)
        k_deposit<true><<<grid, FTHREADS, 0, c->stream>>>(rays, nray, c->H, c->d_ctx, c->d_stats, c->d_ray_len,
                                                          c->d_ray_tm, halt);
    else
        k_deposit<false><<<grid, FTHREADS, 0, c->stream>>>(rays, nray, c->H, c->d_ctx, c->d_stats, c->d_ray_len,
                                                           c->d_ray_tm, halt);
    PSR_HIP(c, hipGetLastError());
    PSR_EV(c, 2, 1);
    return PSR_OK;
}
