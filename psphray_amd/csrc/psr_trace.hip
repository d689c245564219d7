// psr_trace.hip -- per-tick hot path on gfx950: traversal and per-ray sort
// (src/raytrace.c:153-230, src/loop.c:375-386), the tick driver and the
// parity export; the deposition kernel is in psr_deposit.hip.
//
// k_trace   one 256-thread workgroup per ray.  A LIFO stack in LDS holds
//           the first-child index of hit internal nodes; each step expands
//           up to 32 of them (32 x 8 children = 256 box tests), pushing hit
//           internal children and queueing hit leaves; 32 queued leaves
//           (x 8 particle slots) are tested per leaf step.  Every visited box
//           gets the reference's float Pluecker test and every particle the
//           reference's float/double impact test, so the hit set is the
//           reference's exactly.  Hits go to the ray's slot in HBM.
// k_sort    one workgroup per ray: O(n) bucket sort of (d, ipar) in LDS;
//           writes the permutation and the slot / b columns in that order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "psr_device.h"
#include "psr_internal.h"

#define TR_SORT 1024  // lists of up to this many hits are sorted by the trace
// threads per ray of the k_sort tiers: their LDS allows 3 (4096 hits) and 1
// (8192) workgroups per CU, so more threads per ray fill the CU (256^3 sort
// phase: 2.17 ms at 256 and 256, 1.67 at 512 and 512, 3.4 at 1024 and 1024)
#define SORT_NT4 512
#define SORT_NT8 512

// ------------------------------------------------------------ geometry
// segment/AABB test of src/pluecker-sse.c:28-198 (see the oracle for the
// per-pair rule that covers its eight direction classes); float products,
// compiled without contraction
__device__ __forceinline__ bool box_hit(const float s[3], const float dir[3], float len,
                                        const Node &nd) {
    float s2b[3], s2t[3], e2b[3], e2t[3];
    bool sg[3];
#pragma unroll
    for (int i = 0; i < 3; i++) {
        s2b[i] = nd.bot[i] - s[i];
        s2t[i] = nd.top[i] - s[i];
        const float m = len * dir[i];
        e2t[i] = s2t[i] - m;
        e2b[i] = s2b[i] - m;
        sg[i] = dir[i] > 0;
    }
    bool ok = true;
#pragma unroll
    for (int i = 0; i < 3; i++) {
        ok = ok && !(sg[i] ? (s2t[i] < 0) : (s2b[i] > 0));
        ok = ok && !(sg[i] ? (e2b[i] > 0) : (e2t[i] < 0));
    }
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const int i = k == 2 ? 1 : 0, j = k == 0 ? 1 : 2;
        const float A = sg[i] ? s2t[j] : s2b[j];
        const float B = sg[j] ? s2b[i] : s2t[i];
        const float C = sg[i] ? s2b[j] : s2t[j];
        const float D = sg[j] ? s2t[i] : s2b[i];
        const float r0 = dir[i] * A, r1 = dir[j] * B, r2 = dir[i] * C, r3 = dir[j] * D;
        ok = ok && !(r0 < r1 || r2 > r3);
    }
    return ok;
}

// ------------------------------------------------------------ sort
// (d, ipar) order of one ray's hits (the qsort of src/loop.c:383 with ties
// broken by particle index).  Bucket sort in LDS: uniform buckets over
// [dmin, dmax]; the bucket index (int)((d - dmin) * scale) is monotone in d,
// so buckets are ordered and equal distances share a bucket; each bucket
// (about 4 hits) is then insertion-sorted.  O(n) LDS work per ray.
__device__ __forceinline__ bool hit_less(unsigned long long ka, unsigned long long kb, uint32_t ia,
                                         uint32_t ib, const DevHits &H, int64_t base,
                                         const uint32_t *__restrict__ orig) {
    if (ka != kb) return ka < kb;
    const uint32_t pa = orig[H.slot[base + ia]] & ~PSR_INVALID_BIT;
    const uint32_t pb = orig[H.slot[base + ib]] & ~PSR_INVALID_BIT;
    return pa < pb;
}

// the range [min, max] of one value per thread over the NT-thread block
template <int NT>
__device__ __forceinline__ void block_minmax(double &mn, double &mx, double *red) {
    constexpr int NW = NT / 64;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        mn = fmin(mn, __shfl_xor(mn, o, 64));
        mx = fmax(mx, __shfl_xor(mx, o, 64));
    }
    const int wid = threadIdx.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) {
        red[wid] = mn;
        red[NW + wid] = mx;
    }
    __syncthreads();
    mn = red[0];
    mx = red[NW];
#pragma unroll
    for (int k = 1; k < NW; k++) {
        mn = fmin(mn, red[k]);
        mx = fmax(mx, red[NW + k]);
    }
}

template <int MAXN, int MAXB, int NT>
struct SortLds {
    unsigned long long key[MAXN];
    uint16_t idx[MAXN];
    int32_t start[MAXB];
    int32_t fill[MAXB];
    double red[2 * (NT / 64)];
    int32_t wsum[NT / 64];
};

// sorts the n hits of one ray at H[base ..] (n <= MAXN) with the NT
// threads of the block: the permutation and the sorted slot / b columns
template <int MAXN, int MAXB, int NT>
__device__ __forceinline__ void sort_ray(int n, int64_t base, const DevHits &H, const uint32_t *__restrict__ orig,
                                         SortLds<MAXN, MAXB, NT> &L) {
    static_assert(MAXN % NT == 0 && NT % 64 == 0, "whole wavefronts, whole rows of distances");
    constexpr int PER = MAXN / NT;  // distances per thread, held in registers
    const int tid = threadIdx.x;
    // one pass over the list: the range, the bucket counts and the
    // scatter all use these registers
    double dv[PER];
    double mn = INFINITY, mx = -INFINITY;
#pragma unroll
    for (int k = 0; k < PER; k++) {
        const int i = tid + k * NT;
        dv[k] = i < n ? gmem(H.d)[base + i] : 0.0;
        if (i < n) {
            mn = fmin(mn, dv[k]);
            mx = fmax(mx, dv[k]);
        }
    }
    block_minmax<NT>(mn, mx, L.red);
    const double dmin = mn, dmax = mx;
    int B = 1;
    while (B * 4 < n && B < MAXB) B <<= 1;
    const double scale = dmax > dmin ? (double)B / (dmax - dmin) : 0.0;
    for (int b = tid; b < B; b += NT) {
        L.start[b] = 0;
        L.fill[b] = 0;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < PER; k++) {
        if (tid + k * NT < n) {
            const int b = min(B - 1, (int)((dv[k] - dmin) * scale));
            atomicAdd(&L.start[b], 1);
        }
    }
    __syncthreads();
    // exclusive scan of the counts: each thread owns a contiguous run
    {
        const int per = (B + NT - 1) / NT;
        const int b0 = tid * per, b1 = min(B, b0 + per);
        int s = 0;
        for (int b = b0; b < b1; b++) s += L.start[b];
        int incl = s;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const int v = __shfl_up(incl, o, 64);
            if ((tid & 63) >= o) incl += v;
        }
        if ((tid & 63) == 63) L.wsum[tid >> 6] = incl;
        __syncthreads();
        int off = 0;
        for (int w = 0; w < (tid >> 6); w++) off += L.wsum[w];
        int run = off + incl - s;
        for (int b = b0; b < b1; b++) {
            const int c = L.start[b];
            L.start[b] = run;
            run += c;
        }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < PER; k++) {
        const int i = tid + k * NT;
        if (i < n) {
            const int b = min(B - 1, (int)((dv[k] - dmin) * scale));
            const int p = L.start[b] + atomicAdd(&L.fill[b], 1);
            L.key[p] = (unsigned long long)__double_as_longlong(dv[k]);
            L.idx[p] = (uint16_t)i;
        }
    }
    __syncthreads();
    for (int b = tid; b < B; b += NT) {
        const int s0 = L.start[b], e0 = s0 + L.fill[b];
        for (int j = s0 + 1; j < e0; j++) {
            const unsigned long long kj = L.key[j];
            const uint16_t ij = L.idx[j];
            int q = j - 1;
            while (q >= s0 && hit_less(kj, L.key[q], ij, L.idx[q], H, base, orig)) {
                L.key[q + 1] = L.key[q];
                L.idx[q + 1] = L.idx[q];
                q--;
            }
            L.key[q + 1] = kj;
            L.idx[q + 1] = ij;
        }
    }
    __syncthreads();
    // the permutation and the sorted columns: every gather of a thread is
    // issued before its first store (so its PER round trips overlap)
    uint32_t qv[PER], sv[PER];
    double bv[PER];
#pragma unroll
    for (int k = 0; k < PER; k++) {
        const int i = tid + k * NT;
        qv[k] = i < n ? (uint32_t)L.idx[i] : 0u;
    }
#pragma unroll
    for (int k = 0; k < PER; k++) {
        if (tid + k * NT < n) {
            sv[k] = gmem(H.slot)[base + qv[k]];
            bv[k] = gmem(H.b)[base + qv[k]];
        }
    }
#pragma unroll
    for (int k = 0; k < PER; k++) {
        const int i = tid + k * NT;
        if (i < n) {
            H.perm[base + i] = qv[k];
            H.sslot[base + i] = sv[k];
            H.sb[base + i] = bv[k];
        }
    }
    __syncthreads();
}

template <int MAXN, int MAXB, int NT>
__global__ __launch_bounds__(NT) void k_sort(int64_t nray, DevHits H, const uint32_t *__restrict__ orig,
                                             int32_t lo, int32_t hi, const int32_t *halt) {
    __shared__ SortLds<MAXN, MAXB, NT> L;
    if (halted(halt)) return;
    for (int64_t r = blockIdx.x; r < nray; r += gridDim.x) {
        const int n = H.count[r];
        if (n <= lo || n > hi) continue;
        sort_ray<MAXN, MAXB, NT>(n, r * (int64_t)H.cap, H, orig, L);
    }
}

// sorted slot / b columns from a permutation (rays of the global fallback)
__global__ void k_apply_perm(const int32_t *__restrict__ rays_big, int32_t nbig, DevHits H, const int32_t *halt) {
    if (halted(halt)) return;
    for (int b = blockIdx.x; b < nbig; b += gridDim.x) {
        const int64_t r = rays_big[b];
        const int n = H.count[r];
        const int64_t base = r * (int64_t)H.cap;
        for (int i = threadIdx.x; i < n; i += blockDim.x) {
            const uint32_t q = H.perm[base + i];
            H.sslot[base + i] = H.slot[base + q];
            H.sb[base + i] = H.b[base + q];
        }
    }
}

// global-memory bitonic sort for rays beyond the LDS bins (rare): one
// workgroup per listed ray, scratch of P entries per ray
__global__ __launch_bounds__(256) void k_sort_global(const int32_t *__restrict__ rays_big, int32_t nbig,
                                                     DevHits H, const uint32_t *__restrict__ orig,
                                                     unsigned long long *__restrict__ gkey,
                                                     uint32_t *__restrict__ gsec, int32_t pmax,
                                                     const int32_t *halt) {
    if (halted(halt)) return;
    const int tid = threadIdx.x;
    for (int b = blockIdx.x; b < nbig; b += gridDim.x) {
        const int64_t r = rays_big[b];
        const int n = H.count[r];
        const int64_t base = r * (int64_t)H.cap;
        unsigned long long *key = gkey + (int64_t)b * pmax;
        uint32_t *sec = gsec + (int64_t)b * pmax;
        uint32_t *pos = H.perm + base;
        int P = 1;
        while (P < n) P <<= 1;
        // pos needs P entries; the slot holds cap >= n; pad lives in gsec's upper half
        uint32_t *ppos = gsec + (int64_t)nbig * pmax + (int64_t)b * pmax;
        for (int i = tid; i < P; i += 256) {
            if (i < n) {
                key[i] = (unsigned long long)__double_as_longlong(H.d[base + i]);
                sec[i] = orig[H.slot[base + i]] & ~PSR_INVALID_BIT;
            } else {
                key[i] = ~0ull;
                sec[i] = ~0u;
            }
            ppos[i] = (uint32_t)i;
        }
        __syncthreads();
        for (int k = 2; k <= P; k <<= 1) {
            for (int j = k >> 1; j > 0; j >>= 1) {
                for (int i = tid; i < P; i += 256) {
                    const int ixj = i ^ j;
                    if (ixj > i) {
                        const unsigned long long ka = key[i], kb = key[ixj];
                        const uint32_t sa = sec[i], sb = sec[ixj];
                        const bool gt = ka > kb || (ka == kb && sa > sb);
                        const bool up = (i & k) == 0;
                        if (gt == up) {
                            key[i] = kb;
                            key[ixj] = ka;
                            sec[i] = sb;
                            sec[ixj] = sa;
                            const uint32_t t = ppos[i];
                            ppos[i] = ppos[ixj];
                            ppos[ixj] = t;
                        }
                    }
                }
                __threadfence_block();
                __syncthreads();
            }
        }
        for (int i = tid; i < n; i += 256) pos[i] = ppos[i];
        __syncthreads();
    }
}

// total_deposit_count when deposit is skipped (traceOnly)
__global__ void k_count_hits(const int32_t *__restrict__ count, int64_t nray, DevStats *st, const int32_t *halt) {
    if (halted(halt)) return;
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    int v = i < nray ? count[i] : 0;
    v = wave_sum_i(v);
    if ((threadIdx.x & 63) == 0 && v) atomicAdd(&stat_slot(st)->total_deposit_count, (unsigned long long)v);
}

// parity export: CSR lists with caller indices
__global__ void k_export_lists(int64_t nray, DevHits H, const uint32_t *__restrict__ perm,
                               const uint32_t *__restrict__ orig, const int64_t *__restrict__ off,
                               int64_t *__restrict__ oi, double *__restrict__ od, double *__restrict__ ob) {
    const int64_t r = blockIdx.x;
    if (r >= nray) return;
    const int n = H.count[r];
    const int64_t base = r * (int64_t)H.cap;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const uint32_t q = perm[base + i];
        oi[off[r] + i] = (int64_t)(orig[H.slot[base + q]] & ~PSR_INVALID_BIT);
        od[off[r] + i] = H.d[base + q];
        ob[off[r] + i] = H.b[base + q];
    }
}


// NT threads per ray, 256 or 128 (the autotuner times both): a step
// expands NT/8 nodes or tests NT/8 leaves; the LIFO stack holds 16 NT
// entries.  256: 8 waves per SIMD (64 VGPRs; with its sort the trace would
// take 66 and 7); 128: 6 (the sort's distances take 8 registers per thread)
template <int NT>
__global__ __launch_bounds__(NT, NT == 256 ? 8 : 6) void k_trace(
    const psr_ray *__restrict__ rays_arg, const DevTick *__restrict__ tk, int64_t nray,
    const Node *__restrict__ nodes, const int32_t *__restrict__ froot, const float4 *__restrict__ psl,
    const uint32_t *__restrict__ orig, DevHits H, uint8_t *__restrict__ active,
    DevStats *__restrict__ st, const int32_t *halt) {
    // a tick's rays come through its tick record, parity tracing passes them
    const psr_ray *__restrict__ rays = tk ? tk->rays : rays_arg;
    // the traversal's stack and leaf queue, then (same memory) the sort of a
    // list of up to TR_SORT hits
    union TraceLds {
        struct {
            int32_t stack[16 * NT];
            int2 leafq[NT + NT / 4];  // leaves to test: (node, first particle)
        } t;
        SortLds<TR_SORT, TR_SORT / 4, NT> s;
    };
    __shared__ TraceLds U;
    int32_t *const stack = U.t.stack;
    int2 *const leafq = U.t.leafq;
    __shared__ int32_t s_sp, s_lq, s_nh, s_cs, s_cl, s_ovf;
    if (halted(halt)) return;
    const int tid = threadIdx.x;
    for (int64_t r = blockIdx.x; r < nray; r += gridDim.x) {
        const psr_ray R = rays[r];
        const float s[3] = {R.s[0], R.s[1], R.s[2]};
        const float dir[3] = {R.dir[0], R.dir[1], R.dir[2]};
        const float len = (float)R.length;
        const int64_t base = r * (int64_t)H.cap;
        if (tid == 0) {
            s_sp = 0;
            s_lq = 0;
            s_nh = 0;
            s_cs = 0;
            s_cl = 0;
            s_ovf = 0;
            const int root = froot[R.field];
            const Node nd = nodes[root];
            if (nd.b != 0 && box_hit(s, dir, len, nd)) {
                if (nd.b < 0) stack[s_sp++] = nd.a;
                else leafq[s_lq++] = make_int2(root, nd.a);
            }
        }
        for (;;) {
            __syncthreads();
            const int sp = s_sp, lq = s_lq;
            if (sp == 0 && lq == 0) break;
            if (lq >= NT / 8 || sp == 0) {
                const int nl = min(lq, NT / 8);
                const int t = tid >> 3, slot = tid & 7;
                // the leaf's particle range was queued with it, so the
                // particle loads need no node load first
                const int2 L = t < nl ? leafq[lq - nl + t] : make_int2(-1, 0);
                const int k = L.y + slot;
                if (L.x >= 0) {
                    // the leaf's aligned line of eight 16-byte records: an
                    // invalid particle's sml is stored negated (its sign bit
                    // is the flag; slots past the leaf's count are invalid);
                    // the test is computed for it too and its result dropped
                    const float4 p = psl[8 * (int64_t)L.x + slot];
                    // impact test of src/raytrace.c:171-186
                    const float dx = p.x - s[0], dy = p.y - s[1], dz = p.z - s[2];
                    const float px = dx * dir[0], py = dy * dir[1], pz = dz * dir[2];
                    const float qx = dx * dx, qy = dy * dy, qz = dz * dz;
                    double proj = (double)px;
                    proj += (double)py;
                    proj += (double)pz;
                    double d2 = (double)qx;
                    d2 += (double)qy;
                    d2 += (double)qz;
                    const float hh = p.w * p.w;
                    const bool valid = !(__float_as_uint(p.w) >> 31);
                    // The reference's r2 = (sqrt(d2) - proj)(sqrt(d2) + proj)
                    // differs from d2 - proj^2 by under 8 units of 2^-53 d2
                    // (the rounding of the square root, the sum, the
                    // difference and the product), so a particle with hh
                    // below d2 - proj^2 - 2^-48 d2 is outside whatever the
                    // rounding: only the others need the square root and the
                    // reference's expression, which decides for them
                    const double r2a = d2 - proj * proj;
                    if (valid & !((double)hh < r2a - d2 * 3.5527136788005009e-15)) {
                        const double dist = sqrt(d2);
                        const double r2 = (dist - proj) * (dist + proj);
                        if (!((double)hh < r2)) {
                            const int h = atomicAdd(&s_nh, 1);
                            if (h < H.cap) {
                                H.d[base + h] = dist;
                                H.b[base + h] = sqrt(fabs(r2));
                                H.slot[base + h] = (uint32_t)k;
                            }
                            active[k] = 1;
                        }
                    }
                }
                __syncthreads();
                if (tid == 0) s_lq = lq - nl;
            } else {
                const int nn = min(sp, NT / 8);
                const int t = tid >> 3, c = tid & 7;
                const int fc = t < nn ? stack[sp - nn + t] : -1;
                __syncthreads();
                if (fc >= 0) {
                    const int ch = fc + c;
                    const Node C = nodes[ch];
                    if (C.b != 0 && box_hit(s, dir, len, C)) {
                        if (C.b < 0) {
                            const int p = sp - nn + atomicAdd(&s_cs, 1);
                            if (p < 16 * NT) stack[p] = C.a;
                            else s_ovf = 1;
                        } else {
                            const int p = lq + atomicAdd(&s_cl, 1);
                            leafq[p] = make_int2(ch, C.a);
                        }
                    }
                }
                __syncthreads();
                if (tid == 0) {
                    s_sp = min(sp - nn + s_cs, 16 * NT);
                    s_lq = lq + s_cl;
                    s_cs = 0;
                    s_cl = 0;
                }
            }
        }
        const int nh = s_nh;
        if (tid == 0) {
            H.count[r] = nh;
            if (nh > H.cap) atomicOr(&st->flags, PSR_DF_HITCAP);
            if (s_ovf) atomicOr(&st->flags, PSR_DF_STACK);
        }
        // a short list is sorted here, while it is in the cache, by the
        // threads that found it; longer ones by k_sort's tiers
        if (nh > 0 && nh <= TR_SORT) sort_ray<TR_SORT, TR_SORT / 4, NT>(nh, base, H, orig, U.s);
        __syncthreads();
    }
}

// ------------------------------------------------------------ launchers
int psr_ensure_hits(psr_ctx *c, int64_t nray, int32_t cap) {
    if (nray <= c->hit_slots && cap <= c->H.cap && c->H.d) return PSR_OK;
    const int64_t ns = std::max(nray, c->hit_slots);
    const int32_t cp = std::max(cap, c->H.cap);
    // kernels of earlier ticks may still use the old arrays
    PSR_HIP(c, hipStreamSynchronize(c->stream));
    if (c->H.d) (void)hipFree(c->H.d);
    if (c->H.b) (void)hipFree(c->H.b);
    if (c->H.slot) (void)hipFree(c->H.slot);
    if (c->H.count) (void)hipFree(c->H.count);
    if (c->H.perm) (void)hipFree(c->H.perm);
    if (c->H.sslot) (void)hipFree(c->H.sslot);
    if (c->H.sb) (void)hipFree(c->H.sb);
    if (c->H.order) (void)hipFree(c->H.order);
    if (c->d_ray_len) (void)hipFree(c->d_ray_len);
    if (c->d_ray_tm) (void)hipFree(c->d_ray_tm);
    c->H = DevHits{};
    c->d_ray_len = c->d_ray_tm = nullptr;
    c->hit_slots = 0;
    const size_t tot = (size_t)ns * cp;
    PSR_HIP(c, hipMalloc(&c->H.d, sizeof(double) * tot));
    PSR_HIP(c, hipMalloc(&c->H.b, sizeof(double) * tot));
    PSR_HIP(c, hipMalloc(&c->H.slot, sizeof(uint32_t) * tot));
    PSR_HIP(c, hipMalloc(&c->H.perm, sizeof(uint32_t) * tot));
    PSR_HIP(c, hipMalloc(&c->H.sslot, sizeof(uint32_t) * tot));
    PSR_HIP(c, hipMalloc(&c->H.sb, sizeof(double) * tot));
    PSR_HIP(c, hipMalloc(&c->H.order, sizeof(uint32_t) * ns));
    PSR_HIP(c, hipMalloc(&c->H.count, sizeof(int32_t) * ns));
    PSR_HIP(c, hipMalloc(&c->d_ray_len, sizeof(double) * ns));
    PSR_HIP(c, hipMalloc(&c->d_ray_tm, sizeof(double) * ns));
    c->H.cap = cp;
    c->hit_slots = ns;
    return PSR_OK;
}

static int read_flags(psr_ctx *c, int32_t *flags) {
    PSR_HIP(c, hipMemcpyAsync(flags, &c->d_stats->flags, sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
    PSR_HIP(c, hipStreamSynchronize(c->stream));
    return PSR_OK;
}
static int clear_flags(psr_ctx *c) {
    PSR_HIP(c, hipMemsetAsync(&c->d_stats->flags, 0, sizeof(int32_t), c->stream));
    return PSR_OK;
}

static void launch_trace_kernel(psr_ctx *c, int nt, unsigned grid, int64_t ray_first, int64_t nray,
                                const DevTick *tk, const int32_t *halt) {
    if (nt == 128)
        k_trace<128><<<grid, 128, 0, c->stream>>>(c->d_rays + ray_first, tk, nray, c->d_nodes, c->d_field_root,
                                                  c->d_psl, c->d_orig, c->H, c->active_of(c->par), c->d_stats,
                                                  halt);
    else
        k_trace<256><<<grid, 256, 0, c->stream>>>(c->d_rays + ray_first, tk, nray, c->d_nodes, c->d_field_root,
                                                  c->d_psl, c->d_orig, c->H, c->active_of(c->par), c->d_stats,
                                                  halt);
}

int psr_launch_trace(psr_ctx *c, int64_t ray_first, int64_t nray, bool sync, const DevTick *tk, int nt) {
    int32_t cap0 = c->H.cap > 0 ? c->H.cap : 2048;
    int rc = psr_ensure_hits(c, nray, cap0);
    if (rc) return rc;
    const unsigned grid = (unsigned)std::min<int64_t>(nray, 65535 * 4);
    if (!sync) {
        // the flag stays as the previous ticks left it: set only if one of
        // them overflowed, and then this tick is replayed as well
        PSR_EV(c, 0, 0);
        launch_trace_kernel(c, nt, grid, ray_first, nray, tk, &c->d_stats->flags);
        PSR_HIP(c, hipGetLastError());
        PSR_EV(c, 0, 1);
        c->last_nray = nray;
        c->hits_nray = nray;
        return PSR_OK;
    }
    for (int attempt = 0; attempt < 8; attempt++) {
        if ((rc = clear_flags(c))) return rc;
        PSR_EV(c, 0, 0);
        launch_trace_kernel(c, nt, grid, ray_first, nray, tk, nullptr);
        PSR_HIP(c, hipGetLastError());
        PSR_EV(c, 0, 1);
        int32_t fl = 0;
        if ((rc = read_flags(c, &fl))) return rc;
        if (fl & PSR_DF_STACK) PSR_FAIL(c, PSR_ERR_CAPACITY, "traversal stack overflow");
        if (!(fl & PSR_DF_HITCAP)) {
            c->last_nray = nray;
            c->hits_nray = nray;
            return PSR_OK;
        }
        // grow the per-ray hit capacity to the largest list and trace again
        std::vector<int32_t> cnt(nray);
        PSR_HIP(c, hipMemcpy(cnt.data(), c->H.count, sizeof(int32_t) * nray, hipMemcpyDeviceToHost));
        const int32_t mx = *std::max_element(cnt.begin(), cnt.end());
        // capacities stay 2048 * 2^k (a multiple of 64, as the deposit walk needs)
        int32_t ncap = c->H.cap;
        while (ncap < mx) ncap *= 2;
        c->H.cap = 0;  // force reallocation at the new size
        const int64_t ns = c->hit_slots;
        c->hit_slots = 0;
        if ((rc = psr_ensure_hits(c, std::max(ns, nray), ncap))) return rc;
    }
    PSR_FAIL(c, PSR_ERR_CAPACITY, "hit capacity did not converge");
}

int psr_launch_sort(psr_ctx *c, int64_t nray) {
    const unsigned grid = (unsigned)std::min<int64_t>(nray, 65535 * 4);
    const int32_t *halt = &c->d_stats->flags;
    PSR_EV(c, 1, 0);
    // lists of up to TR_SORT hits were sorted by the trace; the longer ones
    // here, in LDS tiers by list length (10 B per hit + 8 B per bucket)
    static_assert(TR_SORT == 1024, "the first tier starts above the trace's sort");
    if (c->sort_split && c->H.cap > 4096) {
        // the 8192-hit tier on stream3 beside the 4096-hit tier: the long
        // lists find room in the CUs the shorter tier's tail leaves idle
        if (!c->stream3) {
            PSR_HIP(c, hipStreamCreateWithFlags(&c->stream3, hipStreamNonBlocking));
            PSR_HIP(c, hipEventCreateWithFlags(&c->ev_tr, hipEventDisableTiming));
            PSR_HIP(c, hipEventCreateWithFlags(&c->ev_s8, hipEventDisableTiming));
        }
        PSR_HIP(c, hipEventRecord(c->ev_tr, c->stream));
        PSR_HIP(c, hipStreamWaitEvent(c->stream3, c->ev_tr, 0));
        k_sort<8192, 2048, SORT_NT8><<<grid, SORT_NT8, 0, c->stream3>>>(nray, c->H, c->d_orig, 4096, 8192, halt);
        PSR_HIP(c, hipGetLastError());
        PSR_HIP(c, hipEventRecord(c->ev_s8, c->stream3));
        k_sort<4096, 1024, SORT_NT4><<<grid, SORT_NT4, 0, c->stream>>>(nray, c->H, c->d_orig, TR_SORT, 4096, halt);
        PSR_HIP(c, hipGetLastError());
        PSR_HIP(c, hipStreamWaitEvent(c->stream, c->ev_s8, 0));
    } else {
        if (c->H.cap > TR_SORT) {
            k_sort<4096, 1024, SORT_NT4><<<grid, SORT_NT4, 0, c->stream>>>(nray, c->H, c->d_orig, TR_SORT, 4096,
                                                                           halt);
            PSR_HIP(c, hipGetLastError());
        }
        if (c->H.cap > 4096) {
            k_sort<8192, 2048, SORT_NT8><<<grid, SORT_NT8, 0, c->stream>>>(nray, c->H, c->d_orig, 4096, 8192,
                                                                           halt);
            PSR_HIP(c, hipGetLastError());
        }
    }
    if (c->H.cap > 8192) {
        std::vector<int32_t> cnt(nray), big;
        int32_t fl = 0;
        PSR_HIP(c, hipMemcpyAsync(cnt.data(), c->H.count, sizeof(int32_t) * nray, hipMemcpyDeviceToHost,
                                  c->stream));
        PSR_HIP(c, hipMemcpyAsync(&fl, &c->d_stats->flags, sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
        PSR_HIP(c, hipStreamSynchronize(c->stream));
        // an overflowed trace leaves counts beyond the capacity: its tick is
        // replayed, nothing to sort now
        if (fl & PSR_DF_HALT) {
            PSR_EV(c, 1, 1);
            return PSR_OK;
        }
        int32_t pmax = 1;
        for (int64_t r = 0; r < nray; r++)
            if (cnt[r] > 8192) {
                big.push_back((int32_t)r);
                while (pmax < cnt[r]) pmax <<= 1;
            }
        if (!big.empty()) {
            const int32_t nb = (int32_t)big.size();
            const size_t need = sizeof(int32_t) * nb + (size_t)nb * pmax * (8 + 4 + 4) + 256;
            int rc = psr_ensure_scratch(c, need);
            if (rc) return rc;
            char *p = (char *)c->d_scratch;
            int32_t *d_big = (int32_t *)p;
            unsigned long long *gkey = (unsigned long long *)(p + ((sizeof(int32_t) * nb + 255) / 256) * 256);
            uint32_t *gsec = (uint32_t *)(gkey + (size_t)nb * pmax);
            PSR_HIP(c, hipMemcpyAsync(d_big, big.data(), sizeof(int32_t) * nb, hipMemcpyHostToDevice, c->stream));
            PSR_HIP(c, hipStreamSynchronize(c->stream));  // 'big' is released on return
            k_sort_global<<<std::min(nb, 1024), 256, 0, c->stream>>>(d_big, nb, c->H, c->d_orig, gkey, gsec,
                                                                     pmax, halt);
            PSR_HIP(c, hipGetLastError());
            k_apply_perm<<<std::min(nb, 1024), 256, 0, c->stream>>>(d_big, nb, c->H, halt);
            PSR_HIP(c, hipGetLastError());
            PSR_HIP(c, hipStreamSynchronize(c->stream));  // scratch is reused by the next call
        }
    }
    PSR_EV(c, 1, 1);
    return PSR_OK;
}

static int check_ready(psr_ctx *c, int64_t ray_first, int64_t nray) {
    if (!c->have_tree) PSR_FAIL(c, PSR_ERR_STATE, "no tree: call psr_tree_build");
    if (!c->have_state) PSR_FAIL(c, PSR_ERR_STATE, "no particle state: call psr_state_upload");
    if (ray_first < 0 || nray < 0 || ray_first + nray > c->nray_pool)
        PSR_FAIL(c, PSR_ERR_ARG, "ray range outside the uploaded pool");
    // a tree rebuilt with fewer fields leaves earlier rays pointing past the roots
    if (nray > 0 && c->ray_field_max >= std::max(c->nfield, 1))
        PSR_FAIL(c, PSR_ERR_ARG, "ray pool names a field the current tree does not have");
    return PSR_OK;
}

extern "C" int psr_trace_lists(psr_ctx *c, int64_t ray_first, int64_t nray, int64_t *offsets, int64_t cap,
                               int64_t *ipar, double *d, double *b) {
    if (!c || !offsets) return PSR_ERR_ARG;
    int rc = check_ready(c, ray_first, nray);
    if (rc) return rc;
    PSR_HIP(c, hipSetDevice(c->device));
    if (int rc_j = psr_join(c)) return rc_j;
    if ((rc = psr_resolve(c))) return rc;
    offsets[0] = 0;
    if (nray == 0) return PSR_OK;
    const int nt = c->trace_req > 0 ? c->trace_req : c->trace_use > 0 ? c->trace_use : 256;
    if ((rc = psr_launch_trace(c, ray_first, nray, true, nullptr, nt))) return rc;
    // parity mode marks particles active like the tick does; clear them
    PSR_HIP(c, hipMemsetAsync(c->active_of(c->par), 0, c->npar > 0 ? c->npar : 1, c->stream));
    if ((rc = psr_launch_sort(c, nray))) return rc;
    std::vector<int32_t> cnt(nray);
    PSR_HIP(c, hipMemcpyAsync(cnt.data(), c->H.count, sizeof(int32_t) * nray, hipMemcpyDeviceToHost, c->stream));
    PSR_HIP(c, hipStreamSynchronize(c->stream));
    for (int64_t r = 0; r < nray; r++) offsets[r + 1] = offsets[r] + cnt[r];
    const int64_t total = offsets[nray];
    if (total > cap || !ipar || !d || !b) {
        c->err = "psr_trace_lists: output capacity smaller than the hit count";
        return PSR_ERR_CAPACITY;
    }
    if (total == 0) return PSR_OK;
    int64_t *d_off = nullptr, *d_i = nullptr;
    double *d_d = nullptr, *d_b = nullptr;
    auto fr = [&]() {
        if (d_off) (void)hipFree(d_off);
        if (d_i) (void)hipFree(d_i);
        if (d_d) (void)hipFree(d_d);
        if (d_b) (void)hipFree(d_b);
    };
    if (hipMalloc(&d_off, sizeof(int64_t) * (nray + 1)) != hipSuccess ||
        hipMalloc(&d_i, sizeof(int64_t) * total) != hipSuccess ||
        hipMalloc(&d_d, sizeof(double) * total) != hipSuccess ||
        hipMalloc(&d_b, sizeof(double) * total) != hipSuccess) {
        fr();
        PSR_FAIL(c, PSR_ERR_HIP, "psr_trace_lists: out of device memory");
    }
    hipError_t e = hipMemcpyAsync(d_off, offsets, sizeof(int64_t) * (nray + 1), hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) {
        k_export_lists<<<(unsigned)nray, 256, 0, c->stream>>>(nray, c->H, c->H.perm, c->d_orig, d_off, d_i, d_d, d_b);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(ipar, d_i, sizeof(int64_t) * total, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(d, d_d, sizeof(double) * total, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(b, d_b, sizeof(double) * total, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    fr();
    if (e != hipSuccess) PSR_FAIL(c, PSR_ERR_HIP, std::string("psr_trace_lists: ") + hipGetErrorString(e));
    return PSR_OK;
}

void psr_tune_reset(psr_ctx *c) {
    c->tune_n = 0;
    c->ttune_n = 0;
    c->trace_use = -1;
    c->tune_file_tried = false;
    c->dep_use = -1;
    c->gtune_n = 0;
    c->graph_use = -1;
}

// the deposit modes the requests leave open, in a fixed order (at most
// DEP_NMODE)
constexpr int DEP_NMODE = 5;
static int dep_candidates(const psr_ctx *c, int *cand) {
    int n = 0;
    const int w = c->walk_req;
    const bool wg = w == PSR_WALK_AUTO || w == PSR_WALK_SINGLE;
    const bool multi = c->accum_req != PSR_ACCUM_BINNED;
    if (wg && c->accum_req != PSR_ACCUM_BINNED) cand[n++] = PSR_DEP_WG_ATOMIC;
    if (wg && c->accum_req != PSR_ACCUM_ATOMIC) cand[n++] = PSR_DEP_WG_BINNED;
    if (multi && (w == PSR_WALK_AUTO || w == PSR_WALK_MULTI)) cand[n++] = PSR_DEP_MULTI;
    if (multi && (w == PSR_WALK_AUTO || w == PSR_WALK_MULTI2)) cand[n++] = PSR_DEP_MULTI2;
    if (multi && (w == PSR_WALK_AUTO || w == PSR_WALK_MULTI8)) cand[n++] = PSR_DEP_MULTI8;
    return n;
}
static_assert(2 * DEP_NMODE <= psr_ctx::DTUNE_N, "two trials per candidate");

// ---- persistence of the decided choices
// One line per workload: "psr-tuning 1 <signature> : dep=<mode> graph=<0|1>",
// the signature naming what the choices depend on (device, particle and
// field counts, physics); -1 for a choice not made.
static std::string tune_signature(const psr_ctx *c) {
    char b[320];
    snprintf(b, sizeof b, "psr-tuning 1 arch=%s cu=%d npar=%lld nfield=%d he=%d iso=%d ots=%d adi=%d",
             c->arch.empty() ? "unknown" : c->arch.c_str(), c->cu_count, (long long)c->npar, (int)c->nfield,
             (int)!c->F.h_only, (int)c->F.isothermal, (int)c->F.on_the_spot, (int)c->F.adiabatic);
    return b;
}

static bool read_lines(const char *path, std::vector<std::string> &lines) {
    FILE *f = fopen(path, "r");
    if (!f) return false;
    char buf[1024];
    while (fgets(buf, sizeof buf, f)) {
        std::string l(buf);
        while (!l.empty() && (l.back() == '\n' || l.back() == '\r')) l.pop_back();
        if (!l.empty()) lines.push_back(l);
    }
    fclose(f);
    return true;
}

static int tune_save(psr_ctx *c, const char *path) {
    const std::string sig = tune_signature(c);
    char tail[64];
    snprintf(tail, sizeof tail, " : dep=%d graph=%d trace=%d", (int)c->dep_use, (int)c->graph_use,
             (int)c->trace_use);
    std::vector<std::string> lines;
    read_lines(path, lines);
    bool put = false;
    for (std::string &l : lines)
        if (l.compare(0, sig.size() + 3, sig + " : ") == 0) {
            l = sig + tail;
            put = true;
        }
    if (!put) lines.push_back(sig + tail);
    const std::string tmp = std::string(path) + ".tmp";
    FILE *f = fopen(tmp.c_str(), "w");
    if (!f) PSR_FAIL(c, PSR_ERR_ARG, "psr_tuning_save: cannot write " + tmp);
    bool ok = true;
    for (const std::string &l : lines) ok = ok && fprintf(f, "%s\n", l.c_str()) > 0;
    if (fclose(f) != 0) ok = false;
    if (!ok || rename(tmp.c_str(), path) != 0) PSR_FAIL(c, PSR_ERR_ARG, std::string("psr_tuning_save: cannot write ") + path);
    return PSR_OK;
}

static int tune_load(psr_ctx *c, const char *path) {
    std::vector<std::string> lines;
    if (!read_lines(path, lines)) PSR_FAIL(c, PSR_ERR_ARG, std::string("psr_tuning_load: cannot read ") + path);
    const std::string key = tune_signature(c) + " : ";
    for (const std::string &l : lines) {
        if (l.compare(0, key.size(), key) != 0) continue;
        int dep = -1, graph = -1, trace = -1;
        if (sscanf(l.c_str() + key.size(), "dep=%d graph=%d trace=%d", &dep, &graph, &trace) < 2)
            PSR_FAIL(c, PSR_ERR_ARG, "psr_tuning_load: malformed line for this workload");
        // a choice applies only where the requests still leave it open
        int cand[8];
        const int nc = dep_candidates(c, cand);
        for (int i = 0; i < nc && nc > 1; i++)
            if (cand[i] == dep) c->dep_use = dep;
        if (c->graph_req < 0 && (graph == 0 || graph == 1)) c->graph_use = graph;
        if (c->trace_req < 0 && (trace == 128 || trace == 256)) c->trace_use = trace;
        return PSR_OK;
    }
    PSR_FAIL(c, PSR_ERR_STATE, std::string("psr_tuning_load: no entry for this workload in ") + path);
}

// PSR_TUNING_FILE: choices of earlier runs applied at this workload's first
// tick, and every new decision written back
static void tune_file_lookup(psr_ctx *c) {
    if (c->tune_file_tried) return;
    c->tune_file_tried = true;
    const char *p = getenv("PSR_TUNING_FILE");
    if (p && *p) {
        const std::string keep = c->err;
        if (tune_load(c, p) != PSR_OK) c->err = keep;
    }
}
static void tune_file_store(psr_ctx *c) {
    const char *p = getenv("PSR_TUNING_FILE");
    if (p && *p) {
        const std::string keep = c->err;
        if (tune_save(c, p) != PSR_OK) c->err = keep;
    }
}

extern "C" int psr_tuning_save(psr_ctx *c, const char *path) {
    if (!c || !path) return PSR_ERR_ARG;
    if (!c->have_tree || !c->have_physics) PSR_FAIL(c, PSR_ERR_STATE, "psr_tuning_save needs a tree and physics");
    return tune_save(c, path);
}

extern "C" int psr_tuning_load(psr_ctx *c, const char *path) {
    if (!c || !path) return PSR_ERR_ARG;
    if (!c->have_tree || !c->have_physics) PSR_FAIL(c, PSR_ERR_STATE, "psr_tuning_load needs a tree and physics");
    if (int rc = psr_resolve(c)) return rc;
    c->tune_file_tried = true;
    return tune_load(c, path);
}

// threads per ray of this tick's trace: the requested or decided one, else
// (a trial tick, *trial its index) the shape traced second; -1 otherwise
static int trace_nt(const psr_ctx *c, int *trial) {
    *trial = -1;
    if (c->trace_req > 0) return c->trace_req;
    if (c->trace_use > 0) return c->trace_use;
    if (c->ttune_n < psr_ctx::TTUNE_N) {
        *trial = c->ttune_n;
        return (c->ttune_n & 1) ? 256 : 128;
    }
    return 256;
}

// the trace trials: each trial tick traces its rays with both shapes, 256
// threads first on even trials and 128 first on odd ones (the second
// launch writes the same hit lists), so both are timed on the same rays;
// the shape with the smaller total wins.  The second launch of a tick finds
// part of the first one's lines still cached, so the decision needs both
// orders: a trial on an overflowed asynchronous tick does not count, and
// the trials start again unless both do.  (Single launches on different
// ticks compared rays of different lengths: 256^3 chose 128 threads in 2 of
// 19 runs that way, trace 2.55 ms instead of 2.30.)
static int ttune_decide(psr_ctx *c, bool block) {
    if (c->trace_req > 0 || c->trace_use > 0 || c->ttune_n < psr_ctx::TTUNE_N) return PSR_OK;
    const hipEvent_t last = c->ttune_ev[4 * psr_ctx::TTUNE_N - 1];
    if (block) PSR_HIP(c, hipEventSynchronize(last));
    else if (hipEventQuery(last) != hipSuccess) return PSR_OK;
    float sum[2] = {0.0f, 0.0f};
    int valid = 0;
    for (int k = 0; k < psr_ctx::TTUNE_N; k++) {
        if (c->ttune_slot[k] >= 0) {
            const int32_t v = ((volatile int32_t *)c->h_ring)[c->ttune_slot[k]];
            if (v == 0 && !block) return PSR_OK;
            if (v == 0 || (v & PSR_DF_HALT)) continue;
        }
        for (int l = 0; l < 2; l++) {
            float ms = 0;
            PSR_HIP(c, hipEventElapsedTime(&ms, c->ttune_ev[4 * k + 2 * l], c->ttune_ev[4 * k + 2 * l + 1]));
            // launch l of trial k: 256 threads for (k + l) even
            sum[(k + l) & 1] += ms;
        }
        valid++;
    }
    if (valid < psr_ctx::TTUNE_N) {
        c->ttune_n = 0;
        return PSR_OK;
    }
    c->trace_use = sum[1] < sum[0] ? 128 : 256;
    tune_file_store(c);
    return PSR_OK;
}

// once every trial has run: keep the mode whose fastest trial deposit was
// fastest.  A trial on an asynchronous tick that overflowed (its kernels
// stopped early, its replay ran later) does not count; trials on replays
// do.  Without a valid trial of each candidate the trials start again.
// Blocks only when `block`.
static int tune_decide(psr_ctx *c, bool block) {
    int cand[DEP_NMODE];
    const int nc = dep_candidates(c, cand);
    if (c->dep_use >= 0 || nc < 2 || c->tune_n < 2 * nc) return PSR_OK;
    const hipEvent_t last = c->tune_ev[2 * (2 * nc) - 1];
    if (block) PSR_HIP(c, hipEventSynchronize(last));
    else if (hipEventQuery(last) != hipSuccess) return PSR_OK;
    // the mean of each candidate's two trials (their ticks are placed
    // symmetrically, see tick_mode)
    float sum[DEP_NMODE], best[DEP_NMODE];
    int cnt[DEP_NMODE];
    for (int m = 0; m < DEP_NMODE; m++) {
        sum[m] = 0.0f;
        cnt[m] = 0;
    }
    for (int k = 0; k < 2 * nc; k++) {
        if (c->tune_slot[k] >= 0) {
            const int32_t v = ((volatile int32_t *)c->h_ring)[c->tune_slot[k]];
            if (v == 0 && !block) return PSR_OK;
            if (v == 0 || (v & PSR_DF_HALT)) continue;
        }
        float ms = 0;
        PSR_HIP(c, hipEventElapsedTime(&ms, c->tune_ev[2 * k], c->tune_ev[2 * k + 1]));
        sum[c->tune_mode[k]] += ms;
        cnt[c->tune_mode[k]]++;
    }
    for (int m = 0; m < DEP_NMODE; m++) best[m] = cnt[m] ? sum[m] / cnt[m] : INFINITY;
    int pick = -1;
    for (int i = 0; i < nc; i++) {
        if (!(best[cand[i]] < INFINITY)) {
            psr_tune_reset(c);
            return PSR_OK;
        }
        if (pick < 0 || best[cand[i]] < best[pick]) pick = cand[i];
    }
    c->dep_use = pick;
    tune_file_store(c);
    return PSR_OK;
}

// the tick record of the kernels that follow (stream order keeps it from
// changing under the previous tick's kernels)
__global__ void k_set_tick(DevTick *tk, const psr_ray *rays, int64_t tick, double tick_time, int32_t *slot) {
    if (threadIdx.x == 0) *tk = DevTick{rays, tick, tick_time, slot, 0};
}

// the flag word after the tick's deposit, into the tick's record: its update
// halts on it and it is what the tick publishes, so an overflow in the next
// tick's trace (which may run beside this tick's update) stops only the
// ticks from that one on
__global__ void k_snap_flags(DevTick *tk, const int32_t *flags) {
    if (threadIdx.x == 0) tk->flags = *(volatile const int32_t *)flags;
}

// the current tick's record and active array: parity p
static void set_parity(psr_ctx *c, int p) {
    c->par = p;
    c->d_tick = c->d_tick_buf + p;
}

__global__ void k_fill_u8(uint8_t *p, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = 0;
}

// the deposit mode of this tick: the only candidate, the decided one, else
// the next trial (candidates in turn); after the last trial and before the
// decision, the first candidate
static int tick_mode(const psr_ctx *c, int *trial) {
    *trial = -1;
    int cand[DEP_NMODE];
    const int nc = dep_candidates(c, cand);
    if (nc == 1) return cand[0];
    if (c->dep_use >= 0) return c->dep_use;
    if (c->tune_n < 2 * nc) {
        // two rounds, the second in reverse order: every candidate's two
        // trial ticks then lie at the same mean position, so a deposit that
        // grows or shrinks steadily over the first ticks (fronts forming,
        // rays cut earlier or later) favours none of them
        *trial = c->tune_n;
        return cand[*trial < nc ? *trial : 2 * nc - 1 - *trial];
    }
    return cand[0];
}

// the tick-graph trials: pairs of consecutive ticks kernel by kernel, then
// as graph launches, twice over; each pair's second tick is timed from the
// end of its first to its own end (the interval between tick ends, which
// counts a kernel-by-kernel tick's update running beside the next tick's
// trace once, not twice), and the faster way is kept.  Pairs broken by
// another tick and trials on ticks that overflowed do not count.
static int gtune_decide(psr_ctx *c, bool block) {
    if (c->graph_use >= 0 || c->gtune_n < psr_ctx::TUNE_N) return PSR_OK;
    const hipEvent_t last = c->gtune_ev[psr_ctx::TUNE_N - 1];
    if (block) PSR_HIP(c, hipEventSynchronize(last));
    else if (hipEventQuery(last) != hipSuccess) return PSR_OK;
    float best[2] = {INFINITY, INFINITY};
    for (int k = 1; k < psr_ctx::TUNE_N; k += 2) {
        if (c->gtune_seq[k] != c->gtune_seq[k - 1] + 1) continue;
        bool ok = true;
        for (int j = k - 1; j <= k; j++) {
            const int32_t v = ((volatile int32_t *)c->h_ring)[c->gtune_slot[j]];
            if (v == 0 && !block) return PSR_OK;
            if (v == 0 || (v & PSR_DF_HALT)) ok = false;
        }
        if (!ok) continue;
        float ms = 0;
        PSR_HIP(c, hipEventElapsedTime(&ms, c->gtune_ev[k - 1], c->gtune_ev[k]));
        best[(k >> 1) & 1] = std::min(best[(k >> 1) & 1], ms);
    }
    if (!(best[0] < INFINITY && best[1] < INFINITY)) {
        c->gtune_n = 0;
        return PSR_OK;
    }
    c->graph_use = best[1] < best[0] ? 1 : 0;
    tune_file_store(c);
    return PSR_OK;
}

// end of a tick: its flag word (with bit 8 set, so 0 means "not yet") into
// the pinned ring slot the host polls
__global__ void k_publish(const DevTick *__restrict__ tk) {
    if (threadIdx.x == 0) *(volatile int32_t *)tk->slot = 0x100 | tk->flags;
}

// the update of the tick on stream2, after the tick's deposit, and its
// publication; the main stream goes on with the next tick's trace and sort
// and waits for this update only before the next deposit (psr_join), which
// reads the particle view the update writes and adds into the sums it clears
static int launch_overlapped_update(psr_ctx *c, const PendingTick &t) {
    int rc;
    // allocations (and the syncs they need) before the fork
    if ((rc = psr_prepare_update(c))) return rc;
    if (!c->stream2) {
        PSR_HIP(c, hipStreamCreateWithFlags(&c->stream2, hipStreamNonBlocking));
        PSR_HIP(c, hipEventCreateWithFlags(&c->ev_dep, hipEventDisableTiming));
        PSR_HIP(c, hipEventCreateWithFlags(&c->ev_upd, hipEventDisableTiming));
    }
    PSR_HIP(c, hipEventRecord(c->ev_dep, c->stream));
    PSR_HIP(c, hipStreamWaitEvent(c->stream2, c->ev_dep, 0));
    const hipStream_t main = c->stream;
    c->stream = c->stream2;
    if (t.chem_mode != PSR_CHEM_EXACT) rc = psr_mlp_update_dev(c, t.chem_mode == PSR_CHEM_SURROGATE_STIFF);
    else rc = psr_launch_update(c);
    if (!rc) {
        k_publish<<<1, 64, 0, c->stream>>>(c->d_tick);
        if (hipGetLastError() != hipSuccess) {
            c->err = "launching the tick's publication";
            rc = PSR_ERR_HIP;
        }
    }
    c->stream = main;
    // recorded even after a failed launch, so that every later call still
    // waits for whatever reached stream2
    const hipError_t e = hipEventRecord(c->ev_upd, c->stream2);
    c->upd_pending = true;
    c->tick_end = c->stream2;
    if (rc) return rc;
    if (e != hipSuccess) PSR_FAIL(c, PSR_ERR_HIP, std::string("recording the update's event: ") + hipGetErrorString(e));
    return PSR_OK;
}

// the kernels of one tick; sync = wait for the trace's capacity check.
// set: write the tick record first (a recorded graph leaves that to its
// launch, since the record is what changes from tick to tick).  publish:
// end with the tick's flag word in its ring slot.  An asynchronous, untimed
// tick launched kernel by kernel runs its update beside the next tick's
// trace and sort (PSR_TUNE_TICK_OVERLAP).
static int tick_body(psr_ctx *c, const PendingTick &t, bool sync, bool set = true, bool publish = false) {
    int rc;
    const int32_t *halt = sync ? nullptr : &c->d_stats->flags;
    const bool overlap = publish && set && !sync && !c->timing && c->overlap_req != 0 &&
                         (t.phases & PSR_PHASE_UPDATE) && c->npar > 0;
    if (set) {
        k_set_tick<<<1, 64, 0, c->stream>>>(c->d_tick, c->d_rays + t.ray_first, t.tick, t.tick_time,
                                            c->h_ring + t.slot);
        PSR_HIP(c, hipGetLastError());
    }
    if (t.nray > 0) {
        int ttrial;
        int nt = trace_nt(c, &ttrial);
        if (sync && ttrial >= 0) {
            // a replay's trace may retrace after growing the hit slots: no
            // trial there
            ttrial = -1;
            nt = 256;
        }
        if (ttrial >= 0) {
            // a timed trace runs alone, first with the other shape
            if ((rc = psr_join(c))) return rc;
            if (!c->ttune_ev[0])
                for (int k = 0; k < 4 * psr_ctx::TTUNE_N; k++) PSR_HIP(c, hipEventCreate(&c->ttune_ev[k]));
            PSR_HIP(c, hipEventRecord(c->ttune_ev[4 * ttrial], c->stream));
            if ((rc = psr_launch_trace(c, t.ray_first, t.nray, sync, c->d_tick, nt == 256 ? 128 : 256))) return rc;
            PSR_HIP(c, hipEventRecord(c->ttune_ev[4 * ttrial + 1], c->stream));
            PSR_HIP(c, hipEventRecord(c->ttune_ev[4 * ttrial + 2], c->stream));
        }
        if ((rc = psr_launch_trace(c, t.ray_first, t.nray, sync, c->d_tick, nt))) return rc;
        if (ttrial >= 0) {
            PSR_HIP(c, hipEventRecord(c->ttune_ev[4 * ttrial + 3], c->stream));
            c->ttune_slot[ttrial] = sync ? -1 : (int32_t)t.slot;
            c->ttune_n++;
        }
        if (t.phases & PSR_PHASE_DEPOSIT) {
            c->sort_split = overlap;
            rc = psr_launch_sort(c, t.nray);
            c->sort_split = false;
            if (rc) return rc;
            int trial;
            const int mode = tick_mode(c, &trial);
            if (mode >= PSR_DEP_MULTI && (rc = psr_launch_ray_order(c, t.nray))) return rc;
            if ((rc = psr_join(c))) return rc;
            if (trial >= 0) {
                // the view build and context upload a deposit may need
                // first stay out of its timing
                if ((rc = psr_prepare_deposit(c, mode))) return rc;
                if (!c->tune_ev[0])
                    for (int k = 0; k < 2 * psr_ctx::DTUNE_N; k++) PSR_HIP(c, hipEventCreate(&c->tune_ev[k]));
                PSR_HIP(c, hipEventRecord(c->tune_ev[2 * trial], c->stream));
            }
            if ((rc = psr_launch_deposit(c, t.ray_first, t.nray, mode))) return rc;
            if (trial >= 0) {
                PSR_HIP(c, hipEventRecord(c->tune_ev[2 * trial + 1], c->stream));
                c->tune_slot[trial] = sync ? -1 : (int32_t)t.slot;
                c->tune_mode[trial] = mode;
                c->tune_n++;
            }
        } else {
            if ((rc = psr_join(c))) return rc;
            k_count_hits<<<(unsigned)((t.nray + 255) / 256), 256, 0, c->stream>>>(c->H.count, t.nray,
                                                                                  c->d_stats, halt);
            PSR_HIP(c, hipGetLastError());
        }
    } else if ((rc = psr_join(c))) {
        return rc;
    }
    k_snap_flags<<<1, 64, 0, c->stream>>>(c->d_tick, &c->d_stats->flags);
    PSR_HIP(c, hipGetLastError());
    c->tick_end = c->stream;
    if (overlap) return launch_overlapped_update(c, t);
    if (t.phases & PSR_PHASE_UPDATE) {
        if (t.chem_mode != PSR_CHEM_EXACT) rc = psr_mlp_update_dev(c, t.chem_mode == PSR_CHEM_SURROGATE_STIFF);
        else rc = psr_launch_update(c);
        if (rc) return rc;
    } else if (c->npar > 0) {
        // merge without update: the active set is dropped (traceOnly); an
        // overflowed tick's replay marks its particles again
        k_fill_u8<<<(unsigned)std::min<int64_t>((c->npar + 255) / 256, 4096), 256, 0, c->stream>>>(
            c->active_of(c->par), c->npar);
        PSR_HIP(c, hipGetLastError());
    }
    if (publish) {
        k_publish<<<1, 64, 0, c->stream>>>(c->d_tick);
        PSR_HIP(c, hipGetLastError());
    }
    return PSR_OK;
}

void psr_graph_reset(psr_ctx *c) {
    if (c->graph_exec) (void)hipGraphExecDestroy(c->graph_exec);
    c->graph_exec = nullptr;
    c->graph_key.clear();
}

// A tick can run as a recorded graph when nothing in it waits on the host:
// no phase timers, no accumulation trial, hit lists within the LDS sort
// tiers (longer lists are sorted through a host-built list).  The shape key
// holds every size and device pointer the tick's launches are given.
static bool graph_usable(const psr_ctx *c, const PendingTick &t) {
    if (c->timing || t.nray <= 0) return false;
    int trial;
    (void)tick_mode(c, &trial);
    if (trial >= 0) return false;
    (void)trace_nt(c, &trial);
    if (trial >= 0) return false;
    return c->H.cap > 0 && c->H.cap <= 8192;
}

static std::vector<int64_t> graph_key(const psr_ctx *c, const PendingTick &t, int mode) {
    const MlpDev M = psr_mlp_dev(c);
    int ttrial;
    return {(int64_t)trace_nt(c, &ttrial),
            t.nray,
            t.phases,
            t.chem_mode,
            (int64_t)mode,
            c->H.cap,
            c->hit_slots,
            (int64_t)(intptr_t)c->H.d,
            (int64_t)(intptr_t)c->d_bin_rec,
            c->bin_n,
            (int64_t)(intptr_t)c->d_qlist,
            (int64_t)(intptr_t)c->d_saved,
            (int64_t)(intptr_t)c->d_q,
            c->npar,
            (int64_t)(intptr_t)c->d_ctx,
            (int64_t)(intptr_t)M.params,
            (int64_t)(intptr_t)M.cap,
            (int64_t)c->mlp_balance,
            (int64_t)(c->mlp_residual * 1e12)};
}

// one tick as a graph launch: allocations first (nothing may allocate or
// wait while recording), the recording when the shape changed, then the
// tick record and the launch
static int tick_graph(psr_ctx *c, const PendingTick &t, bool *done, int *trial) {
    *done = false;
    *trial = -1;
    int rc;
    if (c->graph_req == 0 || (c->graph_req < 0 && c->graph_use == 0)) return PSR_OK;
    const int32_t cap0 = c->H.cap > 0 ? c->H.cap : 2048;
    if ((rc = psr_ensure_hits(c, t.nray, cap0))) return rc;
    if (!graph_usable(c, t)) return PSR_OK;
    if (c->graph_req < 0 && c->graph_use < 0) {
        // a trial tick: two kernel by kernel, two as graph launches, twice
        if (c->gtune_n >= psr_ctx::TUNE_N) return PSR_OK;
        *trial = c->gtune_n;
        if (!c->gtune_ev[0])
            for (int k = 0; k < psr_ctx::TUNE_N; k++) PSR_HIP(c, hipEventCreate(&c->gtune_ev[k]));
        if (!((*trial >> 1) & 1)) return PSR_OK;
    }
    // a graph tick runs after the previous tick's update, on the first tick
    // record and active array (the ones its kernels were recorded with)
    if ((rc = psr_join(c))) return rc;
    set_parity(c, 0);
    int atrial;
    const int mode = tick_mode(c, &atrial);
    if ((rc = psr_sync_dev_ctx(c))) return rc;
    if ((t.phases & PSR_PHASE_DEPOSIT) && (rc = psr_prepare_deposit(c, mode))) return rc;
    if ((t.phases & PSR_PHASE_UPDATE) && (rc = psr_prepare_update(c))) return rc;
    const std::vector<int64_t> key = graph_key(c, t, mode);
    if (!c->graph_exec || key != c->graph_key) {
        psr_graph_reset(c);
        PSR_HIP(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        rc = tick_body(c, t, false, false, true);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(c->stream, &g);
        if (!rc && e != hipSuccess) {
            c->err = std::string("recording the tick graph: ") + hipGetErrorString(e);
            rc = PSR_ERR_HIP;
        }
        if (!rc && hipGraphInstantiate(&c->graph_exec, g, nullptr, nullptr, 0) != hipSuccess) {
            c->err = "instantiating the tick graph failed";
            c->graph_exec = nullptr;
            rc = PSR_ERR_HIP;
        }
        if (g) (void)hipGraphDestroy(g);
        if (rc) return rc;
        c->graph_key = key;
    }
    k_set_tick<<<1, 64, 0, c->stream>>>(c->d_tick, c->d_rays + t.ray_first, t.tick, t.tick_time, c->h_ring + t.slot);
    PSR_HIP(c, hipGetLastError());
    PSR_HIP(c, hipGraphLaunch(c->graph_exec, c->stream));
    c->tick_end = c->stream;
    c->graph_launches++;
    *done = true;
    return PSR_OK;
}

int psr_resolve(psr_ctx *c) {
    if (int rc = psr_join(c)) return rc;
    // no tick pending: every trial tick was published, its events are done
    if (c->pending.empty()) {
        int rc = tune_decide(c, false);
        if (!rc) rc = ttune_decide(c, false);
        return rc ? rc : gtune_decide(c, false);
    }
    PSR_HIP(c, hipStreamSynchronize(c->stream));
    int rc0 = tune_decide(c, true);
    if (!rc0) rc0 = ttune_decide(c, true);
    if (!rc0) rc0 = gtune_decide(c, true);
    if (rc0) return rc0;
    std::deque<PendingTick> todo;
    todo.swap(c->pending);
    // ticks before the first overflow were applied; from it on, every kernel
    // stopped at the halt flag and the state is as that tick found it
    while (!todo.empty()) {
        const int32_t v = ((volatile int32_t *)c->h_ring)[todo.front().slot];
        if (v & PSR_DF_HALT) break;
        todo.pop_front();
    }
    if (todo.empty()) return PSR_OK;
    PSR_HIP(c, hipMemsetAsync(&c->d_stats->flags, 0, sizeof(int32_t), c->stream));
    // a halted trace may have marked part of its particles (in either array):
    // the replays mark them again
    if (c->d_active)
        PSR_HIP(c, hipMemsetAsync(c->d_active, 0, 2 * (size_t)(c->npar > 0 ? c->npar : 1), c->stream));
    for (const PendingTick &t : todo) {
        set_parity(c, (int)(t.slot & 1));
        const int rc = tick_body(c, t, true);
        if (rc) return rc;
    }
    PSR_HIP(c, hipStreamSynchronize(c->stream));
    return tune_decide(c, true);
}

// drop validated ticks from the front of the queue without blocking; an
// overflow found there is replayed at once
static int poll_pending(psr_ctx *c) {
    while (!c->pending.empty()) {
        const int32_t v = ((volatile int32_t *)c->h_ring)[c->pending.front().slot];
        if (v == 0) return PSR_OK;
        if (v & PSR_DF_HALT) return psr_resolve(c);
        c->pending.pop_front();
    }
    return PSR_OK;
}

extern "C" int psr_tick(psr_ctx *c, int64_t ray_first, int64_t nray, int64_t tick, double tick_time,
                        int32_t phases, int32_t chem_mode) {
    if (!c) return PSR_ERR_ARG;
    int rc = check_ready(c, ray_first, nray);
    if (rc) return rc;
    if (!c->have_physics) PSR_FAIL(c, PSR_ERR_STATE, "no physics: call psr_set_physics");
    if (!(phases & PSR_PHASE_TRACE)) PSR_FAIL(c, PSR_ERR_ARG, "a tick starts with the trace phase");
    if (chem_mode != PSR_CHEM_EXACT && chem_mode != PSR_CHEM_SURROGATE && chem_mode != PSR_CHEM_SURROGATE_STIFF)
        PSR_FAIL(c, PSR_ERR_ARG, "unknown chemistry mode");
    PSR_HIP(c, hipSetDevice(c->device));
    if (!c->h_ring) {
        PSR_HIP(c, hipHostMalloc((void **)&c->h_ring, sizeof(int32_t) * PSR_RING, hipHostMallocCoherent));
        memset(c->h_ring, 0, sizeof(int32_t) * PSR_RING);
    }
    if (!c->d_tick_buf) {
        PSR_HIP(c, hipMalloc(&c->d_tick_buf, 2 * sizeof(DevTick)));
        set_parity(c, 0);
    }
    if ((rc = poll_pending(c))) return rc;
    tune_file_lookup(c);
    if ((rc = tune_decide(c, false))) return rc;
    if ((rc = ttune_decide(c, false))) return rc;
    if ((rc = gtune_decide(c, false))) return rc;
    if (c->pending.size() >= PSR_RING - 1 && (rc = psr_resolve(c))) return rc;
    const uint32_t seq = c->seq++;
    PendingTick t{ray_first, nray, tick, tick_time, phases, chem_mode, seq % PSR_RING};
    ((volatile int32_t *)c->h_ring)[t.slot] = 0;
    c->ev_mask = 0;
    PSR_EV(c, 4, 0);
    bool graphed = false;
    int gtrial = -1;
    if ((rc = tick_graph(c, t, &graphed, &gtrial))) return rc;
    if (!graphed) {
        // consecutive ticks alternate records and active arrays
        set_parity(c, (int)(t.slot & 1));
        if ((rc = tick_body(c, t, false, true, true))) return rc;
    }
    if (gtrial >= 0) {
        PSR_HIP(c, hipEventRecord(c->gtune_ev[gtrial], c->tick_end));
        c->gtune_slot[gtrial] = (int32_t)t.slot;
        c->gtune_seq[gtrial] = seq;
        c->gtune_n++;
    }
    c->pending.push_back(t);
    PSR_EV(c, 4, 1);
    if (c->timing) {
        PSR_HIP(c, hipEventSynchronize(c->ev[9]));
        for (int k = 0; k < 5; k++) {
            if (!(c->ev_mask & (1 << k))) continue;
            float a = 0;
            PSR_HIP(c, hipEventElapsedTime(&a, c->ev[2 * k], c->ev[2 * k + 1]));
            c->t_ms[k] += a;
        }
    }
    c->ticks++;
    return PSR_OK;
}

extern "C" int psr_set_timing(psr_ctx *c, int enable) {
    if (!c) return PSR_ERR_ARG;
    PSR_HIP(c, hipSetDevice(c->device));
    if (int rc_j = psr_join(c)) return rc_j;
    if (enable && !c->ev[0])
        for (int k = 0; k < 10; k++) PSR_HIP(c, hipEventCreate(&c->ev[k]));
    c->timing = enable != 0;
    for (int k = 0; k < 5; k++) c->t_ms[k] = 0;
    return PSR_OK;
}

extern "C" int psr_set_tuning(psr_ctx *c, int32_t param, int32_t value) {
    if (!c) return PSR_ERR_ARG;
    if (param == PSR_TUNE_DEPOSIT_ACCUM) {
        if (value != PSR_ACCUM_AUTO && value != PSR_ACCUM_ATOMIC && value != PSR_ACCUM_BINNED)
            PSR_FAIL(c, PSR_ERR_ARG, "unknown deposit accumulation");
        if (value == PSR_ACCUM_BINNED && c->walk_req != PSR_WALK_AUTO && c->walk_req != PSR_WALK_SINGLE)
            PSR_FAIL(c, PSR_ERR_NOTSUP, "several rays per workgroup accumulate with atomics only");
        c->accum_req = value;
        psr_tune_reset(c);
    } else if (param == PSR_TUNE_DEPOSIT_WALK) {
        if (value < PSR_WALK_AUTO || value > PSR_WALK_MULTI8) PSR_FAIL(c, PSR_ERR_ARG, "unknown deposit walk");
        if (value >= PSR_WALK_MULTI && c->accum_req == PSR_ACCUM_BINNED)
            PSR_FAIL(c, PSR_ERR_NOTSUP, "several rays per workgroup accumulate with atomics only");
        c->walk_req = value;
        psr_tune_reset(c);
    } else if (param == PSR_TUNE_TRACE_THREADS) {
        if (value != -1 && value != 128 && value != 256)
            PSR_FAIL(c, PSR_ERR_ARG, "trace threads per ray: -1 (timed trials), 128 or 256");
        c->trace_req = value;
        c->ttune_n = 0;
        c->trace_use = -1;
    } else if (param == PSR_TUNE_TICK_GRAPH) {
        if (value < -1 || value > 1) PSR_FAIL(c, PSR_ERR_ARG, "tick graph: -1 (timed trials), 0 or 1");
        c->graph_req = value;
        c->gtune_n = 0;
        c->graph_use = -1;
    } else if (param == PSR_TUNE_TICK_OVERLAP) {
        if (value != 0 && value != 1) PSR_FAIL(c, PSR_ERR_ARG, "tick overlap: 0 or 1");
        c->overlap_req = value;
        // kernel-by-kernel ticks change speed: the graph trials start again
        c->gtune_n = 0;
        c->graph_use = -1;
    } else {
        PSR_FAIL(c, PSR_ERR_ARG, "unknown tuning parameter");
    }
    return PSR_OK;
}

extern "C" int psr_get_tuning(psr_ctx *c, int32_t param, int32_t *value) {
    if (!c || !value) return PSR_ERR_ARG;
    int cand[DEP_NMODE];
    const int nc = dep_candidates(c, cand);
    const int mode = nc == 1 ? cand[0] : c->dep_use;
    if (param == PSR_TUNE_DEPOSIT_ACCUM)
        *value = c->accum_req != PSR_ACCUM_AUTO ? c->accum_req
                 : mode < 0                     ? PSR_ACCUM_AUTO
                 : mode == PSR_DEP_WG_BINNED    ? PSR_ACCUM_BINNED
                                                : PSR_ACCUM_ATOMIC;
    else if (param == PSR_TUNE_DEPOSIT_WALK)
        *value = c->walk_req != PSR_WALK_AUTO ? c->walk_req
                 : mode < 0                   ? PSR_WALK_AUTO
                 : mode == PSR_DEP_MULTI       ? PSR_WALK_MULTI
                 : mode == PSR_DEP_MULTI2      ? PSR_WALK_MULTI2
                 : mode == PSR_DEP_MULTI8      ? PSR_WALK_MULTI8
                                              : PSR_WALK_SINGLE;
    else if (param == PSR_TUNE_TICK_GRAPH) *value = c->graph_req >= 0 ? c->graph_req : c->graph_use;
    else if (param == PSR_TUNE_TRACE_THREADS) *value = c->trace_req > 0 ? c->trace_req : c->trace_use;
    else if (param == PSR_TUNE_TICK_OVERLAP) *value = c->overlap_req;
    else if (param == PSR_TUNE_GRAPH_LAUNCHES) *value = (int32_t)std::min<int64_t>(c->graph_launches, INT32_MAX);
    else PSR_FAIL(c, PSR_ERR_ARG, "unknown tuning parameter");
    return PSR_OK;
}

extern "C" int psr_get_timers(psr_ctx *c, double *ms) {
    if (!c || !ms) return PSR_ERR_ARG;
    for (int k = 0; k < 5; k++) ms[k] = c->t_ms[k];
    return PSR_OK;
}

extern "C" int psr_ray_results(psr_ctx *c, int64_t nray, double *length, double *tm) {
    if (!c) return PSR_ERR_ARG;
    if (nray > c->last_nray) PSR_FAIL(c, PSR_ERR_ARG, "more rays than the last tick traced");
    PSR_HIP(c, hipSetDevice(c->device));
    if (int rc_j = psr_join(c)) return rc_j;
    if (int rc = psr_resolve(c)) return rc;
    PSR_HIP(c, hipStreamSynchronize(c->stream));
    if (length) PSR_HIP(c, hipMemcpy(length, c->d_ray_len, sizeof(double) * nray, hipMemcpyDeviceToHost));
    if (tm) PSR_HIP(c, hipMemcpy(tm, c->d_ray_tm, sizeof(double) * nray, hipMemcpyDeviceToHost));
    return PSR_OK;
}
