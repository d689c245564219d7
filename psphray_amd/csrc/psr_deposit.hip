// psr_deposit.hip -- photon deposition along the sorted hit lists on gfx950
// (deposit, src/loop.c:388-608).
//
// One wavefront per ray, 64 sorted hits per chunk.  The lanes gather the
// particle data of their hit (the next chunk's gathers are issued before the
// current chunk's serial part), compute the optical depths and expm1, and
// write the values that the transmission recurrence needs to wavefront-
// private LDS, compacted over hits with tausum = 0 (the reference skips
// those without touching TM or testing the cut-off).  The recurrence
//     absorb = min(-TM * expm1(-tau), N_HI) ... ; TM -= absorb
// is then walked serially over LDS broadcast reads (independent of TM, so
// they pipeline), with the cut-off TM / Nph < 1e-10 as one compare against a
// precomputed threshold.  Each lane keeps the TM its hit saw and afterwards
// recomputes its own absorbed photons from it, bit-identically, and scatters
// them with f32 atomics.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "psr_device.h"
#include "psr_internal.h"

namespace {

// one chunk of 64 sorted hits gathered by the lanes of a wavefront
struct DepChunk {
    uint32_t slot;
    double b;
    float sml, mass;
    double xHI, xHeI, xHeII;
    bool valid;
};

template <bool HE>
__device__ __forceinline__ DepChunk dep_load(const DevHits &H, const DevState &S, int64_t base, int i,
                                             int n) {
    DepChunk g;
    g.valid = i < n;
    g.slot = 0;
    g.b = 0;
    g.sml = 1.0f;
    g.mass = 0.0f;
    g.xHI = g.xHeI = g.xHeII = 0.0;
    if (g.valid) {
        g.slot = H.sslot[base + i];
        g.b = H.sb[base + i];
        g.sml = S.ps[g.slot].w;
        g.mass = S.mass[g.slot];
        g.xHI = S.lH[g.slot];
        if (HE) {
            g.xHeI = S.lHeI[g.slot];
            g.xHeII = S.lHeII[g.slot];
        }
    }
    return g;
}

// smallest TM with fl(TM / Nph) >= 1e-10: "TM < 0 || TM / Nph < 1e-10"
// (src/loop.c:585) is then exactly "TM < T0" for Nph > 0, because a
// correctly rounded quotient is monotone in its numerator
__device__ double cutoff_threshold(double Nph) {
    double c = 1e-10 * Nph;
    for (int k = 0; k < 64 && c / Nph >= 1e-10; k++) c = nextafter(c, 0.0);
    for (int k = 0; k < 64 && c / Nph < 1e-10; k++) c = nextafter(c, INFINITY);
    return c;
}

constexpr int DEP_WAVES = 4;

template <bool HE>
struct DepLds {
    double em[DEP_WAVES][64];
    double nhi[DEP_WAVES][64];
    double fhi[HE ? DEP_WAVES : 1][64];
    double fhei[HE ? DEP_WAVES : 1][64];
    double fheii[HE ? DEP_WAVES : 1][64];
    double nhei[HE ? DEP_WAVES : 1][64];
    double nheii[HE ? DEP_WAVES : 1][64];
    uint8_t lane[DEP_WAVES][64];
};

template <bool HE>
__global__ __launch_bounds__(256) void k_deposit(const psr_ray *__restrict__ rays, int64_t nray, DevHits H,
                                                 const DevCtx *__restrict__ C, DevStats *__restrict__ st0,
                                                 double *__restrict__ ray_len, double *__restrict__ ray_tm) {
    __shared__ DepLds<HE> L;
    const DevState &S = C->S;
    const DevTables &T = C->T;
    const DevPhys &P = C->P;
    const int lane = threadIdx.x & 63;
    const int w = threadIdx.x >> 6;
    const unsigned long long lt_mask = (1ull << lane) - 1;
    const int64_t wave0 = blockIdx.x * (int64_t)DEP_WAVES + w;
    const int64_t nwaves = gridDim.x * (int64_t)DEP_WAVES;
    for (int64_t r = wave0; r < nray; r += nwaves) {
        const psr_ray R = rays[r];
        const int n = H.count[r];
        if (lane == 0) {
            ray_len[r] = R.length;
            ray_tm[r] = R.Nph;
        }
        if (n == 0 || R.Nph == 0) continue;
        const int64_t base = r * (int64_t)H.cap;
        double TM = R.Nph;
        if (P.no_photon && R.type < 0) TM = 0.0;
        const double Nph = R.Nph;
        const double T0 = Nph > 0 ? cutoff_threshold(Nph) : 0.0;
        const double freq = R.freq;
        const double sigmaHI = xs_get(T, 0, freq);
        const double sigmaHeI = HE ? xs_get(T, 1, freq) : 0.0;
        const double sigmaHeII = HE ? xs_get(T, 2, freq) : 0.0;
        const double dEHI = fdim(freq, P.C_HI_ENERGY);
        const double dEHeI = fdim(freq, P.C_HEI_ENERGY);
        const double dEHeII = fdim(freq, P.C_HEII_ENERGY);
        const double hf_HI = P.C_H_PER_MASS * dEHI;
        const double hf_HeI = P.C_HE_PER_MASS * dEHeI;
        const double hf_HeII = P.C_HE_PER_MASS * dEHeII;
        const bool secion_terms = P.secondary_ionization && (dEHI > 0 || dEHeI > 0) && T.si_n1 > 0;
        int last = n - 1;
        bool done = false;
        int nsat = 0;
        double fiHI = 0, fiHeI = 0, fiHeII = 0, siHI = 0, siHeI = 0;
        DepChunk g = dep_load<HE>(H, S, base, lane, n);
        for (int c0 = 0; c0 < n && !done; c0 += 64) {
            const DepChunk gn = dep_load<HE>(H, S, base, c0 + 64 + lane, n);
            const uint32_t slot = g.slot;
            const double xHI = g.xHI, xHeI = g.xHeI, xHeII = g.xHeII;
            // src/loop.c:427-477
            const double sml_inv = 1.0 / (double)g.sml;
            const double NH = (double)g.mass * P.C_H_PER_MASS;
            const double NHe = (double)g.mass * P.C_HE_PER_MASS;
            const double NH_inv = 1.0 / NH;
            const double NHe_inv = 1.0 / NHe;
            const double kernel = sph_depth(T, g.b * sml_inv) * (sml_inv * sml_inv) * P.scaling_fac2_inv;
            const double NHI = xHI * NH;
            const double tauHI = sigmaHI * (kernel * NHI);
            double tausum = tauHI;
            double tauHeI = 0, tauHeII = 0;
            const bool he = HE && NHe != 0.0;
            if (he) {
                tauHeI = sigmaHeI * (kernel * (xHeI * NHe));
                tausum += tauHeI;
                tauHeII = sigmaHeII * (kernel * (xHeII * NHe));
                tausum += tauHeII;
            }
            const bool live = g.valid && tausum != 0.0;
            const double em = live ? expm1(-tausum) : 0.0;
            const double NHeI = xHeI * NHe, NHeII = xHeII * NHe;
            double fHI = 1.0, fHeI = 0.0, fHeII = 0.0;
            if (HE && live) {
                fHI = tauHI / tausum;
                fHeI = he ? tauHeI / tausum : 0.0;
                fHeII = he ? tauHeII / tausum : 0.0;
            }
            // compact the live hits into this wavefront's LDS rows
            const unsigned long long lm = __ballot(live);
            const int cnt = __popcll(lm);
            const int p = __popcll(lm & lt_mask);
            if (live) {
                L.em[w][p] = em;
                L.nhi[w][p] = NHI;
                L.lane[w][p] = (uint8_t)lane;
                if (HE) {
                    L.fhi[w][p] = fHI;
                    L.fhei[w][p] = he ? fHeI : 0.0;
                    L.fheii[w][p] = he ? fHeII : 0.0;
                    L.nhei[w][p] = he ? NHeI : -1.0;  // -1 marks "no helium" for the walk
                    L.nheii[w][p] = NHeII;
                }
            }
            __builtin_amdgcn_wave_barrier();
            // transmission recurrence, src/loop.c:479-504, 580-589; for
            // hydrogen only tauHI / tausum is exactly 1
            int cut = -1;
            double my_tm = 0.0;
#pragma unroll 4
            for (int q = 0; q < cnt; q++) {
                const double ae = -TM * L.em[w][q];
                double a = HE ? L.fhi[w][q] * ae : ae;
                const double nh = L.nhi[w][q];
                a = (a > nh) ? nh : a;
                double absorb = a;
                if (HE) {
                    const double n1 = L.nhei[w][q];
                    if (n1 >= 0.0) {
                        double a1 = L.fhei[w][q] * ae;
                        a1 = (a1 > n1) ? n1 : a1;
                        double a2 = L.fheii[w][q] * ae;
                        const double n2 = L.nheii[w][q];
                        a2 = (a2 > n2) ? n2 : a2;
                        absorb += (a1 + a2);
                    }
                }
                my_tm = (p == q) ? TM : my_tm;
                TM -= absorb;
                if (TM < T0) {
                    cut = q;
                    break;
                }
            }
            __builtin_amdgcn_wave_barrier();
            const bool mine = live && (cut < 0 || p <= cut);
            int my_sat = 0;
            if (mine) {
                // this hit's share, recomputed from the TM it saw (same
                // operations as the walk), then src/loop.c:486-573
                const double absorb_est = -my_tm * em;
                double aHI = HE ? fHI * absorb_est : absorb_est;
                if (aHI > NHI) {
                    aHI = NHI;
                    my_sat++;
                }
                double aHeI = 0, aHeII = 0;
                if (he) {
                    aHeI = fHeI * absorb_est;
                    if (aHeI > NHeI) {
                        aHeI = NHeI;
                        my_sat++;
                    }
                    aHeII = fHeII * absorb_est;
                    if (aHeII > NHeII) {
                        aHeII = NHeII;
                        my_sat++;
                    }
                }
                const double deltaHI = aHI * NH_inv;
                const double deltaHeI = he ? aHeI * NHe_inv : 0.0;
                const double deltaHeII = he ? aHeII * NHe_inv : 0.0;
                const double ye = (1.0 - xHI) + P.C_HEMF / P.C_HMF * 0.25 * (xHeII + 2 * (1.0 - xHeII - xHeI));
                double x = ye * NH / (NH + NHe);
                if (x > 1.0) x = 1.0;
                if (P.secondary_ionization && x > 0.0) {
                    atomicAdd(&S.yGdepHI[slot], (float)deltaHI);
                    fiHI += deltaHI * NH;
                    if (secion_terms) {
                        const double log10x = log10(x);
                        const double sec =
                            deltaHI * dEHI / P.C_HI_ENERGY * tab2_get(T, T.si_phi_hi, dEHI, log10x) +
                            deltaHeI * dEHeI / P.C_HI_ENERGY * tab2_get(T, T.si_phi_hi, dEHeI, log10x);
                        atomicAdd(&S.yGdepHI[slot], (float)sec);
                        siHI += sec * NH;
                        atomicAdd(&S.heat[slot], (float)(hf_HI * deltaHI * tab2_get(T, T.si_eh, dEHI, log10x)));
                        if (he) {
                            const double sec2 =
                                deltaHI * dEHI / P.C_HEI_ENERGY * tab2_get(T, T.si_phi_hei, dEHI, log10x) +
                                deltaHeI * dEHeI / P.C_HEI_ENERGY * tab2_get(T, T.si_phi_hei, dEHeI, log10x);
                            atomicAdd(&S.yGdepHeI[slot], (float)(deltaHeI + sec2));
                            siHeI += sec2 * NHe;
                            fiHeI += deltaHeI * NHe;
                            atomicAdd(&S.heat[slot],
                                      (float)(hf_HeI * deltaHeI * tab2_get(T, T.si_eh, dEHeI, log10x)));
                            atomicAdd(&S.yGdepHeII[slot], (float)deltaHeII);
                            fiHeII += deltaHeII * NHe;
                        }
                    } else if (he) {
                        atomicAdd(&S.yGdepHeI[slot], (float)deltaHeI);
                        fiHeI += deltaHeI * NHe;
                        atomicAdd(&S.yGdepHeII[slot], (float)deltaHeII);
                        fiHeII += deltaHeII * NHe;
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
            nsat += wave_sum_i(my_sat);
            if (cut >= 0) {
                last = c0 + L.lane[w][cut];
                done = true;
            }
            __builtin_amdgcn_wave_barrier();
            g = gn;
        }
        // ray results and counters
        const double dl = H.d[base + H.perm[base + last]];
        fiHI = wave_sum(fiHI);
        siHI = wave_sum(siHI);
        if (HE) {
            fiHeI = wave_sum(fiHeI);
            fiHeII = wave_sum(fiHeII);
            siHeI = wave_sum(siHeI);
        }
        if (lane == 0) {
            ray_len[r] = dl;
            ray_tm[r] = TM;
            DevStats *st = stat_slot(st0);
            atomicAdd(&st->total_deposit_count, (unsigned long long)n);
            atomicAdd(&st->deposit_processed, (unsigned long long)(last + 1));
            if (nsat) atomicAdd(&st->saturated_deposit_count, (unsigned long long)nsat);
            atomicAdd(&st->lost_photon_count_sum, TM);
            atomicAdd(&st->first_ionization_HI, fiHI);
            atomicAdd(&st->secondary_ionization_HI, siHI);
            if (HE) {
                atomicAdd(&st->first_ionization_HeI, fiHeI);
                atomicAdd(&st->first_ionization_HeII, fiHeII);
                atomicAdd(&st->secondary_ionization_HeI, siHeI);
            }
        }
    }
}

}  // namespace

int psr_launch_deposit(psr_ctx *c, int64_t ray_first, int64_t nray) {
    int rc = psr_sync_dev_ctx(c);
    if (rc) return rc;
    const unsigned grid =
        (unsigned)std::max<int64_t>(1, std::min<int64_t>((nray + DEP_WAVES - 1) / DEP_WAVES, 65535 * 4));
    PSR_EV(c, 2, 0);
    if (!c->F.h_only)
        k_deposit<true><<<grid, 64 * DEP_WAVES, 0, c->stream>>>(c->d_rays + ray_first, nray, c->H, c->d_ctx,
                                                               c->d_stats, c->d_ray_len, c->d_ray_tm);
    else
        k_deposit<false><<<grid, 64 * DEP_WAVES, 0, c->stream>>>(c->d_rays + ray_first, nray, c->H, c->d_ctx,
                                                                c->d_stats, c->d_ray_len, c->d_ray_tm);
    PSR_HIP(c, hipGetLastError());
    PSR_EV(c, 2, 1);
    return PSR_OK;
}
