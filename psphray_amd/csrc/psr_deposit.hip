// psr_deposit.hip -- photon deposition along the sorted hit lists on gfx950
// (deposit, src/loop.c:388-608).
//
// Three kernels per tick, so that the one serial part carries no memory
// latency beyond its own streamed inputs:
//
//   k_dep_prep   one block per ray, one thread per hit: gathers the particle
//                (sml, mass, x_HI[, x_HeI, x_HeII]) of each sorted hit,
//                computes the optical depths, em = expm1(-tau_sum) and N_HI,
//                and writes them to per-hit arrays.  em = +0 marks a hit with
//                tau_sum = 0, which the reference skips (src/loop.c:475-477);
//                a live hit has em <= -0 (sign bit set).
//   k_dep_walk   one wavefront per ray: the transmission recurrence
//                    TM -= min(-TM * em, N_HI)            (hydrogen only)
//                walked serially, 64 hits per chunk, the chunk's inputs taken
//                from the lanes with readlane (the next chunks' loads are in
//                flight meanwhile).  For hydrogen only the step is evaluated
//                as TM' = max(TM + TM * em, TM - N_HI): correctly rounded
//                subtraction is monotone in the subtrahend, so this is the
//                reference's value bit for bit, with three dependent
//                operations instead of five.  Each lane keeps the TM its hit
//                saw (written to tmb) and the TM after it; the cut-off
//                "TM < 0 || TM / Nph < 1e-10" (src/loop.c:585) is one compare
//                against the threshold T0, tested once per chunk (TM never
//                increases) and resolved to the first live hit by a ballot.
//   k_dep_apply  one block per ray, one thread per hit up to the walk's end:
//                recomputes the hit's absorbed photons from tmb exactly as the
//                walk did, and scatters the ionization / heating deposits with
//                f32 atomics (src/loop.c:486-573).
//
// Keeping the atomics out of the walk matters on gfx9: stores and atomics
// share the in-order vmcnt counter with loads, so a wave that waits for its
// next chunk's loads also waits for every earlier atomic to reach memory.
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

constexpr int DEP_THREADS = 256;

// ---------------------------------------------------------------- prep
// src/loop.c:427-477 per hit
template <bool HE>
__global__ __launch_bounds__(DEP_THREADS) void k_dep_prep(const psr_ray *__restrict__ rays, int64_t nray,
                                                          DevHits H, DevDep D, const DevCtx *__restrict__ C) {
    if (halted(D.halt)) return;
    const DevState &S = C->S;
    const DevTables &T = C->T;
    const DevPhys &P = C->P;
    for (int64_t r = blockIdx.x; r < nray; r += gridDim.x) {
        const int n = H.count[r];
        const double Nph = rays[r].Nph;
        if (n == 0 || Nph == 0) continue;
        const double freq = rays[r].freq;
        const double sigmaHI = xs_get(T, 0, freq);
        const double sigmaHeI = HE ? xs_get(T, 1, freq) : 0.0;
        const double sigmaHeII = HE ? xs_get(T, 2, freq) : 0.0;
        const int64_t base = r * (int64_t)H.cap;
        // no-op entries up to the next multiple of 64 (inside the capacity)
        const int npad = (n + 63) & ~63;
        for (int k = n + threadIdx.x; k < npad; k += DEP_THREADS) {
            const int64_t q = base + k;
            D.em[q] = 0.0;
            D.nhi[q] = 0.0;
            if (HE) {
                D.fhi[q] = 0.0;
                D.fhei[q] = 0.0;
                D.fheii[q] = 0.0;
                D.nhei[q] = -1.0;
                D.nheii[q] = 0.0;
            }
        }
        for (int k = threadIdx.x; k < n; k += DEP_THREADS) {
            const int64_t q = base + k;
            const uint32_t slot = H.sslot[q];
            const double b = H.sb[q];
            const double sml_inv = 1.0 / (double)S.ps[slot].w;
            const double mass = (double)S.mass[slot];
            const double xHI = S.lH[slot];
            const double NH = mass * P.C_H_PER_MASS;
            const double NHe = mass * P.C_HE_PER_MASS;
            const double kernel = sph_depth(T, b * sml_inv) * (sml_inv * sml_inv) * P.scaling_fac2_inv;
            const double NHI = xHI * NH;
            const double tauHI = sigmaHI * (kernel * NHI);
            double tausum = tauHI;
            double tauHeI = 0, tauHeII = 0, NHeI = 0, NHeII = 0;
            const bool he = HE && NHe != 0.0;
            if (he) {
                NHeI = S.lHeI[slot] * NHe;
                NHeII = S.lHeII[slot] * NHe;
                tauHeI = sigmaHeI * (kernel * NHeI);
                tausum += tauHeI;
                tauHeII = sigmaHeII * (kernel * NHeII);
                tausum += tauHeII;
            }
            const bool live = tausum != 0.0;
            // expm1 of a negative argument is <= -0, so the sign bit marks live
            D.em[q] = live ? expm1(-tausum) : 0.0;
            D.nhi[q] = live ? NHI : 0.0;
            if (HE) {
                D.fhi[q] = live ? tauHI / tausum : 0.0;
                D.fhei[q] = (live && he) ? tauHeI / tausum : 0.0;
                D.fheii[q] = (live && he) ? tauHeII / tausum : 0.0;
                D.nhei[q] = (live && he) ? NHeI : -1.0;  // -1: no helium terms
                D.nheii[q] = (live && he) ? NHeII : 0.0;
            }
        }
    }
}

// ---------------------------------------------------------------- walk
template <bool HE>
struct WalkChunk {
    double em, nhi;
    double fhi, fhei, fheii, nhei, nheii;
};

// plain loads, no branch or select next to them (either makes the compiler
// wait for the load at once); the index is clamped into the ray's slots, and
// a chunk that is walked lies below the padded list end, where prep wrote
// no-op entries past n
template <bool HE>
__device__ __forceinline__ WalkChunk<HE> walk_load(const DevDep &D, int64_t base, int k, int cap) {
    WalkChunk<HE> w;
    const int64_t q = base + min(k, cap - 1);
    w.em = D.em[q];
    w.nhi = D.nhi[q];
    if (HE) {
        w.fhi = D.fhi[q];
        w.fhei = D.fhei[q];
        w.fheii = D.fheii[q];
        w.nhei = D.nhei[q];
        w.nheii = D.nheii[q];
    } else {
        w.fhi = w.fhei = w.fheii = w.nheii = 0.0;
        w.nhei = -1.0;
    }
    return w;
}

// block sum of (count, walked, lost) into this block's counter copy
__device__ __forceinline__ void walk_stats(DevStats *st0, unsigned long long cnt, unsigned long long walked,
                                           double lost) {
    __shared__ unsigned long long pc[DEP_THREADS / 64], pw[DEP_THREADS / 64];
    __shared__ double pl[DEP_THREADS / 64];
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        pc[w] = cnt;
        pw[w] = walked;
        pl[w] = lost;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long c = 0, k = 0;
        double l = 0;
        for (int i = 0; i < DEP_THREADS / 64; i++) {
            c += pc[i];
            k += pw[i];
            l += pl[i];
        }
        if (c) {
            DevStats *st = stat_slot(st0);
            atomicAdd(&st->total_deposit_count, c);
            atomicAdd(&st->deposit_processed, k);
            atomicAdd(&st->lost_photon_count_sum, l);
        }
    }
}

// LDS record of one hit for the walk
template <bool HE>
struct WalkRec {
    double em, nhi;
};
template <>
struct WalkRec<true> {
    double em, nhi, fhi, fhei, fheii, nhei, nheii, pad;
};

// a segment of SEG hits, 64 per register set, lane l holding hit q * 64 + l
template <bool HE>
struct WalkSeg {
    static constexpr int K = HE ? 2 : 4;
    static constexpr int SEG = 64 * K;
    WalkChunk<HE> c[K];
};

template <bool HE>
__device__ __forceinline__ void seg_load(WalkSeg<HE> &s, const DevDep &D, int64_t base, int k0, int lane, int cap) {
#pragma unroll
    for (int q = 0; q < WalkSeg<HE>::K; q++) s.c[q] = walk_load<HE>(D, base, k0 + q * 64 + lane, cap);
}

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
__device__ __forceinline__ void put_rec(WalkRec<HE> &r, const WalkChunk<HE> &c) {
    r.em = c.em;
    r.nhi = c.nhi;
    if constexpr (HE) {
        r.fhi = c.fhi;
        r.fhei = c.fhei;
        r.fheii = c.fheii;
        r.nhei = c.nhei;
        r.nheii = c.nheii;
    }
}

template <bool HE>
__global__ __launch_bounds__(DEP_THREADS) void k_dep_walk(const psr_ray *__restrict__ rays, int64_t nray,
                                                          DevHits H, DevDep D, const DevCtx *__restrict__ C,
                                                          DevStats *__restrict__ st0, double *__restrict__ ray_len,
                                                          double *__restrict__ ray_tm) {
    using Seg = WalkSeg<HE>;
    constexpr int SEG = Seg::SEG;
    constexpr int WAVES = DEP_THREADS / 64;
    __shared__ WalkRec<HE> rec[WAVES][SEG];
    __shared__ double tml[WAVES][SEG + 1];
    if (halted(D.halt)) return;
    const int lane = threadIdx.x & 63;
    const int w = threadIdx.x >> 6;
    const int wave0 = __builtin_amdgcn_readfirstlane((int)(blockIdx.x * WAVES + w));
    const int nwaves = gridDim.x * WAVES;
    const int no_photon = C->P.no_photon;
    // this wave's totals, kept by lane 0
    unsigned long long cnt = 0, walked = 0;
    double lost = 0.0;
    for (int64_t r = wave0; r < nray; r += nwaves) {
        const int n = __builtin_amdgcn_readfirstlane(H.count[r]);
        const double Nph = rays[r].Nph;
        if (n == 0 || Nph == 0) {
            if (lane == 0) {
                ray_len[r] = rays[r].length;
                ray_tm[r] = Nph;
                D.last[r] = -1;
            }
            continue;
        }
        const int64_t base = r * (int64_t)H.cap;
        double TM = (no_photon && rays[r].type < 0) ? 0.0 : Nph;
        const double T0 = cutoff_threshold(Nph);
        int last = n - 1;
        double *tmb = D.tmb + base;
        const int cap = H.cap;
        const int npad = (n + 63) & ~63;
        Seg cur, nxt;
        seg_load<HE>(cur, D, base, 0, lane, cap);
        bool cut = false;
        for (int s0 = 0; s0 < npad && !cut; s0 += SEG) {
            // stage the segment in LDS, then start loading the next one
#pragma unroll
            for (int q = 0; q < Seg::K; q++) put_rec<HE>(rec[w][q * 64 + lane], cur.c[q]);
            seg_load<HE>(nxt, D, base, s0 + SEG, lane, cap);
            __builtin_amdgcn_wave_barrier();
            // the serial recurrence on one lane: an LDS broadcast read, the
            // step, and an LDS write of the TM the hit saw
            const int m = min(SEG, npad - s0);
            if (lane == 0) {
                double t = TM;
#pragma unroll 8
                for (int j = 0; j < m; j++) {
                    const WalkRec<HE> v = rec[w][j];
                    tml[w][j] = t;
                    t = walk_step<HE>(t, v);
                }
                tml[w][m] = t;
            }
            __builtin_amdgcn_wave_barrier();
            TM = tml[w][m];
            // TM after hit j is TM before hit j + 1 and never increases: the
            // cut is the first live hit after which TM < T0
#pragma unroll
            for (int q = 0; q < Seg::K; q++) {
                if (q * 64 >= m) break;
                const int c0 = s0 + q * 64;
                const double tb = tml[w][q * 64 + lane];
                const double ta = tml[w][q * 64 + lane + 1];
                // the hit capacity is a multiple of 64 and c0 < n <= capacity,
                // so the chunk lies inside this ray's slots
                tmb[c0 + lane] = tb;
                const unsigned long long cm = __ballot(ta < T0 && c0 + lane < n && is_live(cur.c[q].em));
                if (cm) {
                    const int j = __ffsll((long long)cm) - 1;
                    last = c0 + j;
                    TM = readlane_d(ta, j);
                    cut = true;
                    break;
                }
            }
            __builtin_amdgcn_wave_barrier();
            cur = nxt;
        }
        if (lane == 0) {
            ray_len[r] = H.d[base + H.perm[base + last]];
            ray_tm[r] = TM;
            D.last[r] = last;
            cnt += (unsigned long long)n;
            walked += (unsigned long long)(last + 1);
            lost += TM;
        }
    }
    walk_stats(st0, cnt, walked, lost);
}

// ---------------------------------------------------------------- apply
struct ApplySums {
    int nsat;
    double fiHI, fiHeI, fiHeII, siHI, siHeI;
};

template <bool HE>
__global__ __launch_bounds__(DEP_THREADS) void k_dep_apply(const psr_ray *__restrict__ rays, int64_t nray,
                                                           DevHits H, DevDep D, const DevCtx *__restrict__ C,
                                                           DevStats *__restrict__ st0) {
    const DevState &S = C->S;
    const DevTables &T = C->T;
    const DevPhys &P = C->P;
    if (halted(D.halt)) return;
    ApplySums acc{0, 0, 0, 0, 0, 0};
    for (int64_t r = blockIdx.x; r < nray; r += gridDim.x) {
        const int last = D.last[r];
        if (last < 0) continue;
        const double freq = rays[r].freq;
        const double dEHI = fdim(freq, P.C_HI_ENERGY);
        const double dEHeI = fdim(freq, P.C_HEI_ENERGY);
        const double dEHeII = fdim(freq, P.C_HEII_ENERGY);
        const double hf_HI = P.C_H_PER_MASS * dEHI;
        const double hf_HeI = P.C_HE_PER_MASS * dEHeI;
        const double hf_HeII = P.C_HE_PER_MASS * dEHeII;
        const bool secion_terms = P.secondary_ionization && (dEHI > 0 || dEHeI > 0) && T.si_n1 > 0;
        const int64_t base = r * (int64_t)H.cap;
        for (int k = threadIdx.x; k <= last; k += DEP_THREADS) {
            const int64_t q = base + k;
            const double em = D.em[q];
            if (!is_live(em)) continue;
            const uint32_t slot = H.sslot[q];
            const double NHI = D.nhi[q];
            const double tm = D.tmb[q];
            const double mass = (double)S.mass[slot];
            const double NH = mass * P.C_H_PER_MASS;
            const double NHe = mass * P.C_HE_PER_MASS;
            const double NH_inv = 1.0 / NH;
            const double NHe_inv = 1.0 / NHe;
            // the walk's operations on the TM this hit saw
            const double absorb_est = -tm * em;
            double fHI = 1.0, fHeI = 0.0, fHeII = 0.0, NHeI = 0.0, NHeII = 0.0;
            bool he = false;
            if (HE) {
                fHI = D.fhi[q];
                NHeI = D.nhei[q];
                he = NHeI >= 0.0;
                if (he) {
                    fHeI = D.fhei[q];
                    fHeII = D.fheii[q];
                    NHeII = D.nheii[q];
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
                const double xHI = S.lH[slot];
                const double xHeI = HE ? S.lHeI[slot] : 0.0;
                const double xHeII = HE ? S.lHeII[slot] : 0.0;
                const double ye = (1.0 - xHI) + P.C_HEMF / P.C_HMF * 0.25 * (xHeII + 2 * (1.0 - xHeII - xHeI));
                x = ye * NH / (NH + NHe);
                if (x > 1.0) x = 1.0;
            }
            if (P.secondary_ionization && x > 0.0) {
                atomicAdd(&S.yGdepHI[slot], (float)deltaHI);
                acc.fiHI += deltaHI * NH;
                if (secion_terms) {
                    const double log10x = log10(x);
                    const double sec = deltaHI * dEHI / P.C_HI_ENERGY * tab2_get(T, T.si_phi_hi, dEHI, log10x) +
                                       deltaHeI * dEHeI / P.C_HI_ENERGY * tab2_get(T, T.si_phi_hi, dEHeI, log10x);
                    atomicAdd(&S.yGdepHI[slot], (float)sec);
                    acc.siHI += sec * NH;
                    atomicAdd(&S.heat[slot], (float)(hf_HI * deltaHI * tab2_get(T, T.si_eh, dEHI, log10x)));
                    if (he) {
                        const double sec2 =
                            deltaHI * dEHI / P.C_HEI_ENERGY * tab2_get(T, T.si_phi_hei, dEHI, log10x) +
                            deltaHeI * dEHeI / P.C_HEI_ENERGY * tab2_get(T, T.si_phi_hei, dEHeI, log10x);
                        atomicAdd(&S.yGdepHeI[slot], (float)(deltaHeI + sec2));
                        acc.siHeI += sec2 * NHe;
                        acc.fiHeI += deltaHeI * NHe;
                        atomicAdd(&S.heat[slot], (float)(hf_HeI * deltaHeI * tab2_get(T, T.si_eh, dEHeI, log10x)));
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
                if (hf_HI != 0.0) atomicAdd(&S.heat[slot], (float)(hf_HI * deltaHI));
                if (he) {
                    atomicAdd(&S.yGdepHeI[slot], (float)deltaHeI);
                    atomicAdd(&S.yGdepHeII[slot], (float)deltaHeII);
                    if (hf_HeI != 0.0) atomicAdd(&S.heat[slot], (float)(deltaHeI * hf_HeI));
                    if (hf_HeII != 0.0) atomicAdd(&S.heat[slot], (float)(deltaHeII * hf_HeII));
                }
            }
            atomicAdd(&S.hits[slot], 1);
        }
    }
    // block totals, one atomic per counter and block
    __shared__ double part[DEP_THREADS / 64][6];
    const int w = threadIdx.x >> 6;
    const double ns = (double)wave_sum_i(acc.nsat);
    const double v1 = wave_sum(acc.fiHI), v2 = wave_sum(acc.siHI);
    double v3 = 0, v4 = 0, v5 = 0;
    if (HE) {
        v3 = wave_sum(acc.fiHeI);
        v4 = wave_sum(acc.fiHeII);
        v5 = wave_sum(acc.siHeI);
    }
    if ((threadIdx.x & 63) == 0) {
        part[w][0] = ns;
        part[w][1] = v1;
        part[w][2] = v2;
        part[w][3] = v3;
        part[w][4] = v4;
        part[w][5] = v5;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double t[6] = {0, 0, 0, 0, 0, 0};
        for (int i = 0; i < DEP_THREADS / 64; i++)
            for (int k = 0; k < 6; k++) t[k] += part[i][k];
        DevStats *st = stat_slot(st0);
        if (t[0] != 0) atomicAdd(&st->saturated_deposit_count, (unsigned long long)t[0]);
        if (t[1] != 0) atomicAdd(&st->first_ionization_HI, t[1]);
        if (t[2] != 0) atomicAdd(&st->secondary_ionization_HI, t[2]);
        if (HE) {
            if (t[3] != 0) atomicAdd(&st->first_ionization_HeI, t[3]);
            if (t[4] != 0) atomicAdd(&st->first_ionization_HeII, t[4]);
            if (t[5] != 0) atomicAdd(&st->secondary_ionization_HeI, t[5]);
        }
    }
}

}  // namespace

static int ensure_dep(psr_ctx *c, bool he, int64_t nray) {
    const int64_t slots = c->hit_slots * (int64_t)c->H.cap;
    const int32_t words = he ? 8 : 3;
    if (c->H.cap % 64) PSR_FAIL(c, PSR_ERR_STATE, "deposit: hit capacity must be a multiple of 64");
    if (slots > c->dep_slots || words > c->dep_words || nray > c->dep_rays) {
        // kernels of earlier ticks may still use the old arrays
        PSR_HIP(c, hipStreamSynchronize(c->stream));
    }
    if (slots > c->dep_slots || words > c->dep_words) {
        if (c->d_dep) PSR_HIP(c, hipFree(c->d_dep));
        c->d_dep = nullptr;
        c->dep_slots = 0;
        c->dep_words = 0;
        PSR_HIP(c, hipMalloc(&c->d_dep, sizeof(double) * (size_t)slots * words));
        c->dep_slots = slots;
        c->dep_words = words;
    }
    if (nray > c->dep_rays) {
        if (c->D.last) PSR_HIP(c, hipFree(c->D.last));
        c->D.last = nullptr;
        c->dep_rays = 0;
        PSR_HIP(c, hipMalloc(&c->D.last, sizeof(int32_t) * std::max<int64_t>(nray, c->hit_slots)));
        c->dep_rays = std::max<int64_t>(nray, c->hit_slots);
    }
    c->D.halt = &c->d_stats->flags;
    double *p = c->d_dep;
    const size_t s = (size_t)c->dep_slots;
    c->D.em = p;
    c->D.nhi = p + s;
    c->D.tmb = p + 2 * s;
    if (c->dep_words >= 8) {
        c->D.fhi = p + 3 * s;
        c->D.fhei = p + 4 * s;
        c->D.fheii = p + 5 * s;
        c->D.nhei = p + 6 * s;
        c->D.nheii = p + 7 * s;
    } else {
        c->D.fhi = c->D.fhei = c->D.fheii = c->D.nhei = c->D.nheii = nullptr;
    }
    return PSR_OK;
}

int psr_launch_deposit(psr_ctx *c, int64_t ray_first, int64_t nray) {
    int rc = psr_sync_dev_ctx(c);
    if (rc) return rc;
    const bool he = !c->F.h_only;
    if ((rc = ensure_dep(c, he, nray))) return rc;
    const psr_ray *rays = c->d_rays + ray_first;
    const unsigned gray = (unsigned)std::max<int64_t>(1, std::min<int64_t>(nray, 65535 * 4));
    const unsigned gwalk =
        (unsigned)std::max<int64_t>(1, std::min<int64_t>((nray + DEP_THREADS / 64 - 1) / (DEP_THREADS / 64), 65535 * 4));
    PSR_EV(c, 2, 0);
    if (he) {
        k_dep_prep<true><<<gray, DEP_THREADS, 0, c->stream>>>(rays, nray, c->H, c->D, c->d_ctx);
        k_dep_walk<true><<<gwalk, DEP_THREADS, 0, c->stream>>>(rays, nray, c->H, c->D, c->d_ctx, c->d_stats,
                                                               c->d_ray_len, c->d_ray_tm);
        k_dep_apply<true><<<gray, DEP_THREADS, 0, c->stream>>>(rays, nray, c->H, c->D, c->d_ctx, c->d_stats);
    } else {
        k_dep_prep<false><<<gray, DEP_THREADS, 0, c->stream>>>(rays, nray, c->H, c->D, c->d_ctx);
        k_dep_walk<false><<<gwalk, DEP_THREADS, 0, c->stream>>>(rays, nray, c->H, c->D, c->d_ctx, c->d_stats,
                                                                c->d_ray_len, c->d_ray_tm);
        k_dep_apply<false><<<gray, DEP_THREADS, 0, c->stream>>>(rays, nray, c->H, c->D, c->d_ctx, c->d_stats);
    }
    PSR_HIP(c, hipGetLastError());
    PSR_EV(c, 2, 1);
    return PSR_OK;
}
