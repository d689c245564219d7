// psr_emit.hip -- emit_rays (src/loop.c:163-373) on the device: each tick's
// source rays written straight into the ray pool.
//
// The streams are counter-based (Philox4x32-10 keyed by the seed, blocks
// addressed by (ray id, tick, block)), so the list is a function of the
// seed, the tick and each source's lastemit, the same list a host emitter
// or the oracle builds from the same keys:
//   k_emit_cdf      per emission group (all sources, or one field's), the
//                   running sum of the source weights Ngamma_dot * (tick -
//                   lastemit) * tick_time (src/sources.c:67-78), sequentially
//   k_emit_draw     per ray: the source (the first cumulative weight / total
//                   above a uniform, as the reference's discrete draw picks
//                   with probability weight / total, src/loop.c:174-190), the
//                   isotropic direction (gsl_ran_dir_3d's rejection method,
//                   src/loop.c:250) and the spectrum bin (src/spectra.c:50-60)
//   radix sort      rays by (group, source), stably: the reference's qsort
//                   by (type, isrc) (src/loop.c:236)
//   k_emit_write    the rays in that order; a source's rays share its
//                   weight (src/loop.c:305-321)
//   k_emit_lastemit lastemit = tick for every source that emitted
//                   (src/loop.c:335-338)
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>
#include <vector>

#include "psr_internal.h"

namespace {

constexpr uint32_t EMIT_KEY1 = 0x50535248u;
constexpr int DIR_BLOCKS = 32;  // blocks of direction tries (two each) before (0, 0, 1)

struct Philox4 {
    uint32_t v[4];
};

__device__ __forceinline__ Philox4 philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0,
                                                 uint32_t k1) {
    for (int r = 0; r < 10; r++) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        c0 = hi1 ^ c1 ^ k0;
        c1 = lo1;
        c2 = hi0 ^ c3 ^ k1;
        c3 = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    return Philox4{{c0, c1, c2, c3}};
}

__device__ __forceinline__ double uniform01(uint32_t x) { return ((double)x + 0.5) * (1.0 / 4294967296.0); }

struct EmitSrc {
    double pos[3];
    double ngd;
    int32_t field, spec;
};

struct EmitGroup {
    int64_t src0, nsrc;  // the group's sources: gsrc[src0 .. src0 + nsrc)
    int64_t ray0;        // its first ray id (the stream counter)
    int64_t out0, nray;  // its rays in the output
};

// running sums of the group's source weights; a single thread per group
// adds in source order (the host emitter's cumulative sum)
__global__ void k_emit_cdf(const EmitGroup *__restrict__ grp, int ngroup, const int64_t *__restrict__ gsrc,
                           const EmitSrc *__restrict__ src, const int64_t *__restrict__ lastemit, int64_t tick,
                           double tick_time, double *__restrict__ cdf) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= ngroup) return;
    double run = 0.0;
    for (int64_t k = 0; k < grp[g].nsrc; k++) {
        const int64_t s = gsrc[grp[g].src0 + k];
        run += src[s].ngd * (double)(tick - lastemit[s]) * tick_time;
        cdf[grp[g].src0 + k] = run;
    }
}

__global__ void k_emit_draw(const EmitGroup *__restrict__ grp, int ngroup, int64_t nout,
                            const int64_t *__restrict__ gsrc, const EmitSrc *__restrict__ src,
                            const double *__restrict__ cdf, const int64_t *__restrict__ spec_off,
                            const double *__restrict__ spec_cdf, int64_t tick, uint32_t seed, int sbits,
                            unsigned long long *__restrict__ key, uint32_t *__restrict__ idx,
                            float *__restrict__ dir, double *__restrict__ freq, uint32_t *__restrict__ cnt) {
    const int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (q >= nout) return;
    int g = 0;
    while (g + 1 < ngroup && grp[g + 1].out0 <= q) g++;
    const EmitGroup G = grp[g];
    const uint32_t id = (uint32_t)(G.ray0 + (q - G.out0));
    const uint32_t tlo = (uint32_t)tick, thi = (uint32_t)((uint64_t)tick >> 32);
    const Philox4 b0 = philox4x32_10(id, tlo, thi, 0u, seed, EMIT_KEY1);
    // source: number of cumulative weights / total <= u (at most nsrc - 1)
    const double *c = cdf + G.src0;
    const double total = c[G.nsrc - 1];
    const double u = uniform01(b0.v[0]);
    int64_t lo = 0, hi = G.nsrc;  // first index with c[i] / total > u
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if (c[mid] / total <= u) lo = mid + 1;
        else hi = mid;
    }
    const int64_t pick = lo < G.nsrc - 1 ? lo : G.nsrc - 1;
    const int64_t s = gsrc[G.src0 + pick];
    key[q] = ((unsigned long long)(uint32_t)g << sbits) | (unsigned long long)(uint32_t)s;
    idx[q] = (uint32_t)q;
    atomicAdd(&cnt[s], 1u);
    // direction: x, y uniform in [-1, 1] until x^2 + y^2 <= 1
    double x = 0.0, y = 0.0, z = 1.0;
    bool done = false;
    for (int blk = 1; blk <= DIR_BLOCKS && !done; blk++) {
        const Philox4 w = philox4x32_10(id, tlo, thi, (uint32_t)blk, seed, EMIT_KEY1);
        for (int h = 0; h < 2 && !done; h++) {
            const double xx = -1.0 + 2.0 * uniform01(w.v[2 * h]);
            const double yy = -1.0 + 2.0 * uniform01(w.v[2 * h + 1]);
            const double ss = xx * xx + yy * yy;
            if (ss <= 1.0) {
                const double sc = 2.0 * sqrt(1.0 - ss);
                x = xx * sc;
                y = yy * sc;
                z = -1.0 + 2.0 * ss;
                done = true;
            }
        }
    }
    dir[3 * q + 0] = (float)x;
    dir[3 * q + 1] = (float)y;
    dir[3 * q + 2] = (float)z;
    // spectrum bin: number of cumulative bin weights <= u (at most nbin - 1)
    const int32_t sp = src[s].spec;
    const int64_t a = spec_off[sp], nb = spec_off[sp + 1] - a;
    const double u3 = uniform01(b0.v[3]);
    int64_t l2 = 0, h2 = nb;
    while (l2 < h2) {
        const int64_t mid = (l2 + h2) >> 1;
        if (spec_cdf[a + mid] <= u3) l2 = mid + 1;
        else h2 = mid;
    }
    freq[q] = (double)(l2 < nb - 1 ? l2 : nb - 1);  // bin index; k_emit_write looks up its energy
}

__global__ void k_emit_write(int64_t nout, int sbits, const unsigned long long *__restrict__ skey,
                             const uint32_t *__restrict__ sidx, const EmitSrc *__restrict__ src,
                             const int64_t *__restrict__ lastemit, const uint32_t *__restrict__ cnt,
                             const float *__restrict__ dir, const double *__restrict__ fbin,
                             const int64_t *__restrict__ spec_off, const double *__restrict__ spec_eng,
                             const double *__restrict__ boxsizes, int64_t tick, double tick_time,
                             psr_ray *__restrict__ out) {
    const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (p >= nout) return;
    const int64_t s = (int64_t)(skey[p] & ((1ull << sbits) - 1));
    const uint32_t q = sidx[p];
    const EmitSrc S = src[s];
    psr_ray r;
    for (int d = 0; d < 3; d++) {
        r.s[d] = (float)S.pos[d];
        r.dir[d] = dir[3 * (int64_t)q + d];
    }
    const double w = S.ngd * (double)(tick - lastemit[s]) * tick_time;
    r.Nph = w * (1.0 / (double)cnt[s]);
    r.freq = spec_eng[spec_off[S.spec] + (int64_t)fbin[q]];
    r.length = boxsizes[S.field] * 2;
    r.type = -1;
    r.field = S.field;
    r.src_or_par = s;
    out[p] = r;
}

__global__ void k_emit_lastemit(int64_t nsrc, uint32_t *__restrict__ cnt, int64_t *__restrict__ lastemit,
                                int64_t tick) {
    const int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (s >= nsrc) return;
    if (cnt[s]) lastemit[s] = tick;
    cnt[s] = 0;
}

__global__ void k_emit_clear(int64_t nsrc, uint32_t *__restrict__ cnt) {
    const int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (s < nsrc) cnt[s] = 0;
}

// ---- recombination rays (src/loop.c:197-232, :273-299, :339-368)
// The candidates are marked in caller order (marker[caller] = tree slot + 1),
// compacted in ascending caller index, and each species' draw weights
// yGrec_s / x_HI summed in chunks of REC_CHUNK candidates: within a chunk
// sequentially, the chunk totals sequentially, each candidate's cumulative
// weight = the chunks before it + its running sum in its chunk (the oracle
// sums in the same order, so both pick the same particles).
constexpr int REC_CHUNK = 1024;

__global__ void k_rec_mark(DevHits H, int64_t nr, const uint32_t *__restrict__ perm, uint32_t *__restrict__ marker) {
    for (int64_t r = blockIdx.x; r < nr; r += gridDim.x) {
        const int n = min(H.count[r], H.cap);
        const uint32_t *sl = H.slot + r * (int64_t)H.cap;
        for (int j = threadIdx.x; j < n; j += blockDim.x) {
            const uint32_t k = sl[j];
            marker[perm[k]] = k + 1;
        }
    }
}

__global__ void k_rec_flag(int64_t n, const uint32_t *__restrict__ marker, int32_t *__restrict__ flag) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n) flag[i] = marker[i] != 0;
}

__global__ void k_rec_compact(int64_t n, const uint32_t *__restrict__ marker, const int32_t *__restrict__ pos,
                              uint32_t *__restrict__ cand_slot, uint32_t *__restrict__ cand_ipar) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n || !marker[i]) return;
    cand_slot[pos[i]] = marker[i] - 1;
    cand_ipar[pos[i]] = (uint32_t)i;
}

// running sums of the weights within each chunk; chunk totals into ctot
__global__ void k_rec_chunk(int64_t ncand, const uint32_t *__restrict__ cand_slot, const float *__restrict__ yg,
                            const double *__restrict__ lH, double *__restrict__ cdf, double *__restrict__ ctot) {
    const int64_t b = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    const int64_t j0 = b * REC_CHUNK;
    if (j0 >= ncand) return;
    const int64_t j1 = min(ncand, j0 + REC_CHUNK);
    double run = 0.0;
    for (int64_t j = j0; j < j1; j++) {
        const uint32_t k = cand_slot[j];
        double x = lH[k];
        if (x == 0) x = 1e-10;
        run += (double)yg[k] / x;
        cdf[j] = run;
    }
    ctot[b] = run;
}

// chunk offsets (exclusive running sum, in place) and the total at [nchunk]
__global__ void k_rec_offsets(int64_t nchunk, double *__restrict__ ctot) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    double run = 0.0;
    for (int64_t b = 0; b < nchunk; b++) {
        const double t = ctot[b];
        ctot[b] = run;
        run += t;
    }
    ctot[nchunk] = run;
}

__global__ void k_rec_cdf(int64_t ncand, double *__restrict__ cdf, const double *__restrict__ ctot) {
    const int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (j < ncand) cdf[j] = ctot[j / REC_CHUNK] + cdf[j];
}

// nrec draws of species s: key (s << 32) | caller index, value the draw
__global__ void k_rec_draw(int s, int64_t nrec, int64_t ncand, const double *__restrict__ cdf,
                           const double *__restrict__ ctot, int64_t nchunk, const uint32_t *__restrict__ cand_ipar,
                           int64_t tick, uint32_t seed, unsigned long long *__restrict__ key,
                           uint32_t *__restrict__ val) {
    const int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (q >= nrec) return;
    const double total = ctot[nchunk];
    const Philox4 b0 = philox4x32_10((uint32_t)q, (uint32_t)tick, (uint32_t)((uint64_t)tick >> 32),
                                     (uint32_t)(s + 1) << 16, seed, EMIT_KEY1);
    const double u = uniform01(b0.v[0]);
    int64_t lo = 0, hi = ncand;  // first candidate with cdf / total > u
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if (cdf[mid] / total <= u) lo = mid + 1;
        else hi = mid;
    }
    const int64_t pick = lo < ncand - 1 ? lo : ncand - 1;
    key[q] = ((unsigned long long)s << 32) | cand_ipar[pick];
    val[q] = (uint32_t)q;
}

// threshold mode: flags of the candidates with yGrec_s > threshold
__global__ void k_rec_above(int64_t ncand, const uint32_t *__restrict__ cand_slot, const float *__restrict__ yg,
                            double threshold, int32_t *__restrict__ flag) {
    const int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (j < ncand) flag[j] = (double)yg[cand_slot[j]] > threshold;
}
__global__ void k_rec_above_keys(int s, int64_t ncand, const int32_t *__restrict__ flag,
                                 const int32_t *__restrict__ rank, const uint32_t *__restrict__ cand_ipar,
                                 unsigned long long *__restrict__ key, uint32_t *__restrict__ val) {
    const int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (j >= ncand || !flag[j]) return;
    key[rank[j]] = ((unsigned long long)s << 32) | cand_ipar[j];
    val[rank[j]] = (uint32_t)rank[j];
}

struct RecTabs {
    const double *eng;    // [3][PSR_LTE_NFREQ]
    const double *cdf;    // [3][PSR_LTE_NLOGT][PSR_LTE_NFREQ]
    const double *tb;     // [PSR_LTE_NLOGT - 1]
};

// the rays, sorted by (species, particle), each group's draws in order
__global__ void k_rec_write(int64_t n, const unsigned long long *__restrict__ skey, const uint32_t *__restrict__ sval,
                            const uint32_t *__restrict__ marker, DevState S, DevPhys P, RecTabs L, int disable_lte,
                            double length, int64_t tick, uint32_t seed, psr_ray *__restrict__ out) {
    const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (p >= n) return;
    const unsigned long long kk = skey[p];
    const int s = (int)(kk >> 32);
    const uint32_t ipar = (uint32_t)kk;
    const uint32_t k = marker[ipar] - 1;
    // the particle's rays of this species: [lower, upper) of its key
    int64_t lo = 0, hi = p;
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if (skey[mid] < kk) lo = mid + 1;
        else hi = mid;
    }
    const int64_t first = lo;
    lo = p + 1;
    hi = n;
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if (skey[mid] <= kk) lo = mid + 1;
        else hi = mid;
    }
    const int64_t cnt = lo - first;
    const float *yg = s == 0 ? S.yGrecHII : s == 1 ? S.yGrecHeII : S.yGrecHeIII;
    const double N = (double)S.mass[k] * (s == 0 ? P.C_H_PER_MASS : P.C_HE_PER_MASS);
    const uint32_t q = sval[p];
    const uint32_t tlo = (uint32_t)tick, thi = (uint32_t)((uint64_t)tick >> 32);
    const uint32_t sb = (uint32_t)(s + 1) << 16;
    const Philox4 b0 = philox4x32_10(q, tlo, thi, sb, seed, EMIT_KEY1);
    psr_ray r;
    const float4 ps = S.ps[k];
    r.s[0] = ps.x;
    r.s[1] = ps.y;
    r.s[2] = ps.z;
    double x = 0.0, y = 0.0, z = 1.0;
    bool done = false;
    for (int blk = 1; blk <= DIR_BLOCKS && !done; blk++) {
        const Philox4 w = philox4x32_10(q, tlo, thi, sb | (uint32_t)blk, seed, EMIT_KEY1);
        for (int h = 0; h < 2 && !done; h++) {
            const double xx = -1.0 + 2.0 * uniform01(w.v[2 * h]);
            const double yy = -1.0 + 2.0 * uniform01(w.v[2 * h + 1]);
            const double ss = xx * xx + yy * yy;
            if (ss <= 1.0) {
                const double sc = 2.0 * sqrt(1.0 - ss);
                x = xx * sc;
                y = yy * sc;
                z = -1.0 + 2.0 * ss;
                done = true;
            }
        }
    }
    r.dir[0] = (float)x;
    r.dir[1] = (float)y;
    r.dir[2] = (float)z;
    r.Nph = ((double)yg[k] * N) * (1.0 / (double)cnt);
    // LTE photon energy at the particle's temperature (lte_gen_freq, src/lte.c:34-45)
    const double *eng = L.eng + (size_t)s * PSR_LTE_NFREQ;
    if (disable_lte) {
        r.freq = eng[0];
    } else {
        const double T = ieye2T(P, S.ie[k], lambda_ye(P, S.lH[k], S.lHeI[k], S.lHeII[k]));
        int row = 0;
        while (row < PSR_LTE_NLOGT - 1 && L.tb[row] <= T) row++;
        const double *c = L.cdf + ((size_t)s * PSR_LTE_NLOGT + row) * PSR_LTE_NFREQ;
        const double u3 = uniform01(b0.v[3]);
        int lo2 = 0, hi2 = PSR_LTE_NFREQ;
        while (lo2 < hi2) {
            const int mid = (lo2 + hi2) >> 1;
            if (c[mid] <= u3) lo2 = mid + 1;
            else hi2 = mid;
        }
        r.freq = eng[lo2 < PSR_LTE_NFREQ - 1 ? lo2 : PSR_LTE_NFREQ - 1];
    }
    r.length = length;
    r.type = s;
    r.field = 0;
    r.src_or_par = ipar;
    out[p] = r;
}

// yGrec_s -= Nph / N_s for each ray, in ray order, clamped at 0
// (src/loop.c:339-368); one thread per particle and species
__global__ void k_rec_debit(int64_t n, const unsigned long long *__restrict__ skey, const uint32_t *__restrict__ marker,
                            DevState S, DevPhys P, const psr_ray *__restrict__ rays) {
    const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (p >= n || (p > 0 && skey[p - 1] == skey[p])) return;
    const unsigned long long kk = skey[p];
    const int s = (int)(kk >> 32);
    const uint32_t k = marker[(uint32_t)kk] - 1;
    float *yg = s == 0 ? S.yGrecHII : s == 1 ? S.yGrecHeII : S.yGrecHeIII;
    const double N = (double)S.mass[k] * (s == 0 ? P.C_H_PER_MASS : P.C_HE_PER_MASS);
    float v = yg[k];
    for (int64_t q = p; q < n && skey[q] == kk; q++) {
        v = (float)((double)v - rays[q].Nph / N);
        if (v < 0) v = 0;
    }
    yg[k] = v;
}

__global__ void k_fill_u32(int64_t n, uint32_t *__restrict__ a) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n) a[i] = 0;
}

inline unsigned blocks(int64_t n, int t = 256) { return (unsigned)std::max<int64_t>(1, (n + t - 1) / t); }

}  // namespace

// ------------------------------------------------------------- host side
struct EmitState {
    std::vector<EmitSrc> hsrc;
    std::vector<int32_t> field;
    int64_t nsrc = 0;
    EmitSrc *d_src = nullptr;
    int64_t *d_lastemit = nullptr;
    uint32_t *d_cnt = nullptr;
    double *d_cdf = nullptr;
    int64_t *d_gsrc = nullptr;  // sources of the groups, by field then index
    std::vector<int64_t> gsrc_off;  // per field: first entry in d_gsrc (nfield + 1)
    int32_t nspec = 0;
    int64_t *d_spec_off = nullptr;
    double *d_spec_eng = nullptr, *d_spec_cdf = nullptr;
    int64_t last_tick = INT64_MIN;
    // per-call buffers
    int64_t cap = 0;
    unsigned long long *d_key = nullptr, *d_skey = nullptr;
    uint32_t *d_idx = nullptr, *d_sidx = nullptr;
    float *d_dir = nullptr;
    double *d_fbin = nullptr;
    EmitGroup *d_grp = nullptr;
    int64_t grp_cap = 0;
    double *d_box = nullptr;
    int64_t box_n = 0;
    // recombination rays: per particle (caller order) the candidate marker,
    // flags and their scan; candidates' tree slots and caller indices and
    // their cumulative weights; chunk totals; the draws' keys and values
    int64_t rec_npar = 0;
    uint32_t *d_marker = nullptr, *d_cslot = nullptr, *d_cipar = nullptr;
    int32_t *d_flag = nullptr, *d_pos = nullptr;
    double *d_ccdf = nullptr, *d_ctot = nullptr;
    int64_t rec_cap = 0;
    unsigned long long *d_rkey = nullptr, *d_rskey = nullptr;
    uint32_t *d_rval = nullptr, *d_rsval = nullptr;
};

template <typename T>
static void efree(T *&p) {
    if (p) (void)hipFree(p);
    p = nullptr;
}

void psr_emit_free(psr_ctx *c) {
    EmitState *e = c->emit;
    if (!e) return;
    efree(e->d_src);
    efree(e->d_lastemit);
    efree(e->d_cnt);
    efree(e->d_cdf);
    efree(e->d_gsrc);
    efree(e->d_spec_off);
    efree(e->d_spec_eng);
    efree(e->d_spec_cdf);
    efree(e->d_key);
    efree(e->d_skey);
    efree(e->d_idx);
    efree(e->d_sidx);
    efree(e->d_dir);
    efree(e->d_fbin);
    efree(e->d_grp);
    efree(e->d_box);
    efree(e->d_marker);
    efree(e->d_cslot);
    efree(e->d_cipar);
    efree(e->d_flag);
    efree(e->d_pos);
    efree(e->d_ccdf);
    efree(e->d_ctot);
    efree(e->d_rkey);
    efree(e->d_rskey);
    efree(e->d_rval);
    efree(e->d_rsval);
    delete e;
    c->emit = nullptr;
}

extern "C" int psr_sources_upload(psr_ctx *c, int64_t nsrc, const double *pos, const double *ngamma_dot,
                                  const int32_t *field, const int32_t *spec, const int64_t *lastemit) {
    if (!c || nsrc < 0 || (nsrc > 0 && (!pos || !ngamma_dot))) return PSR_ERR_ARG;
    if (nsrc >= (int64_t)UINT32_MAX) PSR_FAIL(c, PSR_ERR_ARG, "psr_sources_upload: too many sources");
    PSR_HIP(c, hipSetDevice(c->device));
    if (int rc_j = psr_join(c)) return rc_j;
    if (int rc = psr_resolve(c)) return rc;
    std::vector<EmitSrc> hs(nsrc);
    int32_t nf = 1;
    for (int64_t i = 0; i < nsrc; i++) {
        for (int d = 0; d < 3; d++) hs[i].pos[d] = pos[3 * i + d];
        hs[i].ngd = ngamma_dot[i];
        hs[i].field = field ? field[i] : 0;
        hs[i].spec = spec ? spec[i] : 0;
        if (hs[i].field < 0 || hs[i].field > 65535 || hs[i].spec < 0 || !(hs[i].ngd >= 0))
            PSR_FAIL(c, PSR_ERR_ARG, "psr_sources_upload: field / spectrum index or Ngamma_dot out of range");
        nf = std::max(nf, hs[i].field + 1);
    }
    psr_emit_free(c);
    c->emit = new EmitState();
    EmitState *e = c->emit;
    e->hsrc = hs;
    e->nsrc = nsrc;
    const int64_t na = std::max<int64_t>(nsrc, 1);
    PSR_HIP(c, hipMalloc(&e->d_src, sizeof(EmitSrc) * na));
    PSR_HIP(c, hipMalloc(&e->d_lastemit, sizeof(int64_t) * na));
    PSR_HIP(c, hipMalloc(&e->d_cnt, sizeof(uint32_t) * na));
    PSR_HIP(c, hipMalloc(&e->d_cdf, sizeof(double) * na));
    PSR_HIP(c, hipMalloc(&e->d_gsrc, sizeof(int64_t) * 2 * na));
    std::vector<int64_t> le(na, 0);
    if (lastemit)
        for (int64_t i = 0; i < nsrc; i++) le[i] = lastemit[i];
    // sources of each field, in index order, then every source in index order
    std::vector<int64_t> gs;
    e->gsrc_off.assign(nf + 2, 0);
    for (int32_t f = 0; f < nf; f++) {
        e->gsrc_off[f] = (int64_t)gs.size();
        for (int64_t i = 0; i < nsrc; i++)
            if (hs[i].field == f) gs.push_back(i);
    }
    e->gsrc_off[nf] = (int64_t)gs.size();
    for (int64_t i = 0; i < nsrc; i++) gs.push_back(i);
    e->gsrc_off[nf + 1] = (int64_t)gs.size();
    if (nsrc) {
        PSR_HIP(c, hipMemcpyAsync(e->d_src, hs.data(), sizeof(EmitSrc) * nsrc, hipMemcpyHostToDevice, c->stream));
        PSR_HIP(c, hipMemcpyAsync(e->d_lastemit, le.data(), sizeof(int64_t) * nsrc, hipMemcpyHostToDevice,
                                  c->stream));
        PSR_HIP(c, hipMemcpyAsync(e->d_gsrc, gs.data(), sizeof(int64_t) * gs.size(), hipMemcpyHostToDevice,
                                  c->stream));
    }
    k_emit_clear<<<blocks(na), 256, 0, c->stream>>>(na, e->d_cnt);
    PSR_HIP(c, hipGetLastError());
    PSR_HIP(c, hipStreamSynchronize(c->stream));
    return PSR_OK;
}

extern "C" int psr_spectra_upload(psr_ctx *c, int32_t nspec, const int64_t *offsets, const double *energy,
                                  const double *cdf) {
    if (!c || nspec < 1 || !offsets || !energy || !cdf) return PSR_ERR_ARG;
    if (!c->emit) PSR_FAIL(c, PSR_ERR_STATE, "psr_spectra_upload before psr_sources_upload");
    if (offsets[0] != 0) PSR_FAIL(c, PSR_ERR_ARG, "spectrum offsets must start at 0");
    for (int32_t k = 0; k < nspec; k++) {
        if (offsets[k + 1] <= offsets[k]) PSR_FAIL(c, PSR_ERR_ARG, "every spectrum needs a bin");
        for (int64_t b = offsets[k]; b + 1 < offsets[k + 1]; b++)
            if (!(cdf[b] <= cdf[b + 1])) PSR_FAIL(c, PSR_ERR_ARG, "spectrum cdf must be non-decreasing");
    }
    EmitState *e = c->emit;
    for (const EmitSrc &s : e->hsrc)
        if (s.spec >= nspec) PSR_FAIL(c, PSR_ERR_ARG, "a source names a spectrum beyond nspec");
    PSR_HIP(c, hipSetDevice(c->device));
    if (int rc_j = psr_join(c)) return rc_j;
    efree(e->d_spec_off);
    efree(e->d_spec_eng);
    efree(e->d_spec_cdf);
    const int64_t nb = offsets[nspec];
    PSR_HIP(c, hipMalloc(&e->d_spec_off, sizeof(int64_t) * (nspec + 1)));
    PSR_HIP(c, hipMalloc(&e->d_spec_eng, sizeof(double) * nb));
    PSR_HIP(c, hipMalloc(&e->d_spec_cdf, sizeof(double) * nb));
    PSR_HIP(c, hipMemcpy(e->d_spec_off, offsets, sizeof(int64_t) * (nspec + 1), hipMemcpyHostToDevice));
    PSR_HIP(c, hipMemcpy(e->d_spec_eng, energy, sizeof(double) * nb, hipMemcpyHostToDevice));
    PSR_HIP(c, hipMemcpy(e->d_spec_cdf, cdf, sizeof(double) * nb, hipMemcpyHostToDevice));
    e->nspec = nspec;
    return PSR_OK;
}

extern "C" int psr_sources_lastemit(psr_ctx *c, int64_t *lastemit) {
    if (!c || !lastemit) return PSR_ERR_ARG;
    if (!c->emit) PSR_FAIL(c, PSR_ERR_STATE, "no sources");
    PSR_HIP(c, hipSetDevice(c->device));
    if (int rc_j = psr_join(c)) return rc_j;
    if (int rc = psr_resolve(c)) return rc;
    PSR_HIP(c, hipStreamSynchronize(c->stream));
    if (c->emit->nsrc)
        PSR_HIP(c, hipMemcpy(lastemit, c->emit->d_lastemit, sizeof(int64_t) * c->emit->nsrc, hipMemcpyDeviceToHost));
    return PSR_OK;
}

extern "C" int psr_set_recombination(psr_ctx *c, int64_t nrec, double threshold, int32_t disable_lte) {
    if (!c || !(threshold >= 0)) return PSR_ERR_ARG;
    c->rec_nrec = nrec;
    c->rec_threshold = threshold;
    c->rec_disable_lte = disable_lte ? 1 : 0;
    return PSR_OK;
}

// the recombination draws of a tick, sorted by (species, particle) into
// e->d_rskey / d_rsval; *n their count (0 when none are due)
static int rec_plan(psr_ctx *c, EmitState *e, int64_t tick, uint32_t seed, int64_t cut, int64_t *n) {
    *n = 0;
    if (c->rec_nrec == 0 || c->F.on_the_spot || tick <= 1) return PSR_OK;
    if (!c->have_state) PSR_FAIL(c, PSR_ERR_STATE, "recombination rays need the particle state");
    if (c->nfield > 1) PSR_FAIL(c, PSR_ERR_NOTSUP, "recombination rays on batched fields");
    if (!c->d_lte) PSR_FAIL(c, PSR_ERR_STATE, "recombination rays need psr_set_physics");
    // the hit lists of the last traced tick (an overflowed one replayed first)
    if (int rc = psr_resolve(c)) return rc;
    const int64_t nr = std::min(c->hits_nray, cut);
    const int64_t np = c->npar;
    if (nr <= 0 || np == 0) return PSR_OK;
    if (np > e->rec_npar) {
        efree(e->d_marker);
        efree(e->d_cslot);
        efree(e->d_cipar);
        efree(e->d_flag);
        efree(e->d_pos);
        efree(e->d_ccdf);
        efree(e->d_ctot);
        PSR_HIP(c, hipMalloc(&e->d_marker, sizeof(uint32_t) * np));
        PSR_HIP(c, hipMalloc(&e->d_cslot, sizeof(uint32_t) * np));
        PSR_HIP(c, hipMalloc(&e->d_cipar, sizeof(uint32_t) * np));
        PSR_HIP(c, hipMalloc(&e->d_flag, sizeof(int32_t) * np));
        PSR_HIP(c, hipMalloc(&e->d_pos, sizeof(int32_t) * np));
        PSR_HIP(c, hipMalloc(&e->d_ccdf, sizeof(double) * np));
        PSR_HIP(c, hipMalloc(&e->d_ctot, sizeof(double) * (np / REC_CHUNK + 2)));
        e->rec_npar = np;
    }
    if (psr_ensure_scratch(c, psr_scan_scratch_i32(np))) return PSR_ERR_HIP;
    k_fill_u32<<<blocks(np), 256, 0, c->stream>>>(np, e->d_marker);
    k_rec_mark<<<(unsigned)std::min<int64_t>(nr, 65535 * 4), 256, 0, c->stream>>>(c->H, nr, c->d_perm, e->d_marker);
    k_rec_flag<<<blocks(np), 256, 0, c->stream>>>(np, e->d_marker, e->d_flag);
    PSR_HIP(c, hipGetLastError());
    PSR_HIP(c, psr_exclusive_sum_i32(e->d_flag, e->d_pos, np, c->d_scratch, c->stream));
    k_rec_compact<<<blocks(np), 256, 0, c->stream>>>(np, e->d_marker, e->d_pos, e->d_cslot, e->d_cipar);
    PSR_HIP(c, hipGetLastError());
    int32_t tail[2] = {0, 0};
    PSR_HIP(c, hipMemcpyAsync(&tail[0], e->d_pos + np - 1, sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
    PSR_HIP(c, hipMemcpyAsync(&tail[1], e->d_flag + np - 1, sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
    PSR_HIP(c, hipStreamSynchronize(c->stream));
    const int64_t ncand = (int64_t)tail[0] + tail[1];
    if (ncand == 0) return PSR_OK;
    const int ns = c->F.h_only ? 1 : 3;
    const DevState S = c->state();
    const float *yg[3] = {S.yGrecHII, S.yGrecHeII, S.yGrecHeIII};
    // the draws' counts per species first (threshold mode: a scan of flags)
    int64_t cnt[3] = {0, 0, 0}, total_n = 0;
    const int64_t nchunk = (ncand + REC_CHUNK - 1) / REC_CHUNK;
    for (int sp = 0; sp < ns; sp++) {
        if (c->rec_nrec > 0) {
            k_rec_chunk<<<blocks(nchunk, 64), 64, 0, c->stream>>>(ncand, e->d_cslot, yg[sp], S.lH, e->d_ccdf, e->d_ctot);
            k_rec_offsets<<<1, 1, 0, c->stream>>>(nchunk, e->d_ctot);
            PSR_HIP(c, hipGetLastError());
            double tot = 0;
            PSR_HIP(c, hipMemcpyAsync(&tot, e->d_ctot + nchunk, sizeof(double), hipMemcpyDeviceToHost, c->stream));
            PSR_HIP(c, hipStreamSynchronize(c->stream));
            // no weight (every yGrec_s zero): no rays of this species
            cnt[sp] = tot > 0 ? c->rec_nrec : 0;
        } else {
            k_rec_above<<<blocks(ncand), 256, 0, c->stream>>>(ncand, e->d_cslot, yg[sp], c->rec_threshold, e->d_flag);
            PSR_HIP(c, hipGetLastError());
            PSR_HIP(c, psr_exclusive_sum_i32(e->d_flag, e->d_pos, ncand, c->d_scratch, c->stream));
            PSR_HIP(c, hipMemcpyAsync(&tail[0], e->d_pos + ncand - 1, sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
            PSR_HIP(c, hipMemcpyAsync(&tail[1], e->d_flag + ncand - 1, sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
            PSR_HIP(c, hipStreamSynchronize(c->stream));
            cnt[sp] = (int64_t)tail[0] + tail[1];
        }
        total_n += cnt[sp];
    }
    if (total_n == 0) return PSR_OK;
    if (total_n >= (int64_t)INT32_MAX) PSR_FAIL(c, PSR_ERR_ARG, "psr_emit_rays: too many recombination rays");
    if (total_n > e->rec_cap) {
        efree(e->d_rkey);
        efree(e->d_rskey);
        efree(e->d_rval);
        efree(e->d_rsval);
        PSR_HIP(c, hipMalloc(&e->d_rkey, sizeof(unsigned long long) * total_n));
        PSR_HIP(c, hipMalloc(&e->d_rskey, sizeof(unsigned long long) * total_n));
        PSR_HIP(c, hipMalloc(&e->d_rval, sizeof(uint32_t) * total_n));
        PSR_HIP(c, hipMalloc(&e->d_rsval, sizeof(uint32_t) * total_n));
        e->rec_cap = total_n;
    }
    int64_t off = 0;
    for (int sp = 0; sp < ns; sp++) {
        if (cnt[sp] == 0) continue;
        if (c->rec_nrec > 0) {
            k_rec_chunk<<<blocks(nchunk, 64), 64, 0, c->stream>>>(ncand, e->d_cslot, yg[sp], S.lH, e->d_ccdf, e->d_ctot);
            k_rec_offsets<<<1, 1, 0, c->stream>>>(nchunk, e->d_ctot);
            k_rec_cdf<<<blocks(ncand), 256, 0, c->stream>>>(ncand, e->d_ccdf, e->d_ctot);
            k_rec_draw<<<blocks(cnt[sp]), 256, 0, c->stream>>>(sp, cnt[sp], ncand, e->d_ccdf, e->d_ctot, nchunk,
                                                              e->d_cipar, tick, seed, e->d_rkey + off,
                                                              e->d_rval + off);
        } else {
            k_rec_above<<<blocks(ncand), 256, 0, c->stream>>>(ncand, e->d_cslot, yg[sp], c->rec_threshold, e->d_flag);
            PSR_HIP(c, psr_exclusive_sum_i32(e->d_flag, e->d_pos, ncand, c->d_scratch, c->stream));
            k_rec_above_keys<<<blocks(ncand), 256, 0, c->stream>>>(sp, ncand, e->d_flag, e->d_pos, e->d_cipar,
                                                                   e->d_rkey + off, e->d_rval + off);
        }
        PSR_HIP(c, hipGetLastError());
        off += cnt[sp];
    }
    // (species, particle) keys: 2 + 32 bits, stable in the draw order
    if (psr_ensure_scratch(c, psr_sort_scratch_u64(total_n))) return PSR_ERR_HIP;
    PSR_HIP(c, psr_sort_pairs_u64(e->d_rkey, e->d_rval, e->d_rskey, e->d_rsval, total_n, 0, 34, c->d_scratch,
                                  c->stream));
    *n = total_n;
    return PSR_OK;
}

extern "C" int psr_emit_rays(psr_ctx *c, int64_t tick, double tick_time, int64_t nray, int64_t per_field,
                             uint32_t seed, int64_t pool_first, int64_t *nray_out) {
    if (!c || nray < 0 || per_field < 0 || pool_first < 0) return PSR_ERR_ARG;
    EmitState *e = c->emit;
    if (!e || e->nspec == 0) PSR_FAIL(c, PSR_ERR_STATE, "psr_emit_rays needs psr_sources_upload and psr_spectra_upload");
    if (!c->have_tree) PSR_FAIL(c, PSR_ERR_STATE, "psr_emit_rays needs a tree (the fields' box sizes)");
    if (tick <= e->last_tick)
        PSR_FAIL(c, PSR_ERR_ARG, "psr_emit_rays: ticks are emitted in increasing order, once each");
    for (const EmitSrc &s : e->hsrc)
        if (s.field >= std::max(c->nfield, 1)) PSR_FAIL(c, PSR_ERR_ARG, "a source names a field the tree does not have");
    PSR_HIP(c, hipSetDevice(c->device));
    // recombination rays are drawn from the particle state the last tick's
    // update writes; source rays read none of it, and their kernels may run
    // beside that update (the next tick's trace follows them in stream order)
    if (c->rec_nrec > 0)
        if (int rc_j = psr_join(c)) return rc_j;
    // groups: one of every source, or one per field; a group emits when it
    // has rays and a source with Ngamma_dot > 0 (every lastemit < tick)
    const int32_t nf = (int32_t)e->gsrc_off.size() - 2;
    std::vector<EmitGroup> G;
    int64_t nout = 0;
    std::vector<int64_t> hgs;  // host copy of the group member lists
    {
        std::vector<int64_t> tmp(e->gsrc_off.back());
        // rebuild the member lists on the host (the same order as on the device)
        int64_t w = 0;
        for (int32_t f = 0; f < nf; f++)
            for (int64_t i = 0; i < e->nsrc; i++)
                if (e->hsrc[i].field == f) tmp[w++] = i;
        for (int64_t i = 0; i < e->nsrc; i++) tmp[w++] = i;
        hgs.swap(tmp);
    }
    auto group = [&](int64_t a, int64_t b, int64_t nr, int64_t ray0) {
        if (nr == 0 || b == a) return;
        bool lit = false;
        for (int64_t k = a; k < b; k++) lit = lit || e->hsrc[hgs[k]].ngd > 0;
        if (!lit) return;
        G.push_back(EmitGroup{a, b - a, ray0, nout, nr});
        nout += nr;
    };
    if (per_field == 0) {
        group(e->gsrc_off[nf], e->gsrc_off[nf + 1], nray, 0);
    } else {
        for (int32_t f = 0; f < nf; f++) group(e->gsrc_off[f], e->gsrc_off[f + 1], per_field, f * per_field);
    }
    if ((int64_t)G.size() >= (int64_t)1 << 31 || nout >= (int64_t)UINT32_MAX)
        PSR_FAIL(c, PSR_ERR_ARG, "psr_emit_rays: too many rays");
    // recombination rays from the last traced tick's candidates, after the
    // source rays; the next tick's candidates come from the first nout + 1
    // rays of this list
    const int64_t cut = c->rec_cut;
    int64_t nrec = 0;
    if (int rc = rec_plan(c, e, tick, seed, cut, &nrec)) return rc;
    c->rec_cut = nout + 1;
    if (nray_out) *nray_out = nout + nrec;
    // a pending tick whose rays overlap the range (an overflow replay needs
    // them) is resolved first; so is a pool that must grow
    const int64_t end = pool_first + nout + nrec;
    bool overlap = end > c->ray_cap;
    for (const PendingTick &t : c->pending)
        if (t.ray_first < end && pool_first < t.ray_first + t.nray) overlap = true;
    if (overlap)
        if (int rc = psr_resolve(c)) return rc;
    if (end > c->ray_cap) {
        psr_ray *nr = nullptr;
        PSR_HIP(c, hipMalloc(&nr, sizeof(psr_ray) * std::max<int64_t>(end, 2 * c->ray_cap)));
        if (c->nray_pool)
            PSR_HIP(c, hipMemcpyAsync(nr, c->d_rays, sizeof(psr_ray) * c->nray_pool, hipMemcpyDeviceToDevice, c->stream));
        PSR_HIP(c, hipStreamSynchronize(c->stream));
        if (c->d_rays) PSR_HIP(c, hipFree(c->d_rays));
        c->d_rays = nr;
        c->ray_cap = std::max<int64_t>(end, 2 * c->ray_cap);
    }
    e->last_tick = tick;
    if (nrec) {
        const DevState S = c->state();
        const DevPhys P = psr_dev_phys(c);
        RecTabs L;
        L.eng = c->d_lte;
        L.cdf = c->d_lte + 3 * PSR_LTE_NFREQ;
        L.tb = L.cdf + (size_t)3 * PSR_LTE_NLOGT * PSR_LTE_NFREQ;
        psr_ray *out = c->d_rays + pool_first + nout;
        // the recombination ray length, src/loop.c:169
        const double length = c->boxsizes[0] * 2 * 0.1;
        k_rec_write<<<blocks(nrec), 256, 0, c->stream>>>(nrec, e->d_rskey, e->d_rsval, e->d_marker, S, P, L,
                                                        c->rec_disable_lte, length, tick, seed, out);
        k_rec_debit<<<blocks(nrec), 256, 0, c->stream>>>(nrec, e->d_rskey, e->d_marker, S, P, out);
        PSR_HIP(c, hipGetLastError());
        c->nray_pool = std::max(c->nray_pool, end);
        c->ray_field_max = std::max(c->ray_field_max, 0);
    }
    if (nout == 0) return PSR_OK;
    if (nout > e->cap) {
        efree(e->d_key);
        efree(e->d_skey);
        efree(e->d_idx);
        efree(e->d_sidx);
        efree(e->d_dir);
        efree(e->d_fbin);
        PSR_HIP(c, hipMalloc(&e->d_key, sizeof(unsigned long long) * nout));
        PSR_HIP(c, hipMalloc(&e->d_skey, sizeof(unsigned long long) * nout));
        PSR_HIP(c, hipMalloc(&e->d_idx, sizeof(uint32_t) * nout));
        PSR_HIP(c, hipMalloc(&e->d_sidx, sizeof(uint32_t) * nout));
        PSR_HIP(c, hipMalloc(&e->d_dir, sizeof(float) * 3 * nout));
        PSR_HIP(c, hipMalloc(&e->d_fbin, sizeof(double) * nout));
        e->cap = nout;
    }
    if ((int64_t)G.size() > e->grp_cap) {
        efree(e->d_grp);
        PSR_HIP(c, hipMalloc(&e->d_grp, sizeof(EmitGroup) * G.size()));
        e->grp_cap = (int64_t)G.size();
    }
    if ((int64_t)std::max(c->nfield, 1) > e->box_n) {
        efree(e->d_box);
        PSR_HIP(c, hipMalloc(&e->d_box, sizeof(double) * std::max(c->nfield, 1)));
        e->box_n = std::max(c->nfield, 1);
    }
    PSR_HIP(c, hipMemcpyAsync(e->d_grp, G.data(), sizeof(EmitGroup) * G.size(), hipMemcpyHostToDevice, c->stream));
    PSR_HIP(c, hipMemcpyAsync(e->d_box, c->boxsizes.data(), sizeof(double) * c->boxsizes.size(),
                              hipMemcpyHostToDevice, c->stream));
    const int ng = (int)G.size();
    k_emit_cdf<<<blocks(ng, 64), 64, 0, c->stream>>>(e->d_grp, ng, e->d_gsrc, e->d_src, e->d_lastemit, tick,
                                                     tick_time, e->d_cdf);
    PSR_HIP(c, hipGetLastError());
    // keys (group << sbits) | source, sorted stably over their bits
    int gbits = 0, sbits = 1;
    while (((int64_t)1 << gbits) < ng) gbits++;
    while (((int64_t)1 << sbits) < e->nsrc) sbits++;
    k_emit_draw<<<blocks(nout), 256, 0, c->stream>>>(e->d_grp, ng, nout, e->d_gsrc, e->d_src, e->d_cdf,
                                                     e->d_spec_off, e->d_spec_cdf, tick, seed, sbits, e->d_key,
                                                     e->d_idx, e->d_dir, e->d_fbin, e->d_cnt);
    PSR_HIP(c, hipGetLastError());
    if (psr_ensure_scratch(c, psr_sort_scratch_u64(nout))) return PSR_ERR_HIP;
    PSR_HIP(c, psr_sort_pairs_u64(e->d_key, e->d_idx, e->d_skey, e->d_sidx, nout, 0, sbits + gbits, c->d_scratch,
                                  c->stream));
    k_emit_write<<<blocks(nout), 256, 0, c->stream>>>(nout, sbits, e->d_skey, e->d_sidx, e->d_src, e->d_lastemit, e->d_cnt,
                                                      e->d_dir, e->d_fbin, e->d_spec_off, e->d_spec_eng, e->d_box,
                                                      tick, tick_time, c->d_rays + pool_first);
    PSR_HIP(c, hipGetLastError());
    k_emit_lastemit<<<blocks(e->nsrc), 256, 0, c->stream>>>(e->nsrc, e->d_cnt, e->d_lastemit, tick);
    PSR_HIP(c, hipGetLastError());
    c->nray_pool = std::max(c->nray_pool, end);
    int32_t fmax = 0;
    for (const EmitSrc &sr : e->hsrc) fmax = std::max(fmax, sr.field);
    c->ray_field_max = std::max(c->ray_field_max, fmax);
    return PSR_OK;
}
