// psr_api.hip -- context, epoch set-up (device octree build) and the C ABI.
//
// Octree build (replaces rt_switch_epoch, src/raytrace.c:41-151).  The
// reference inserts particles one by one, splitting a cell when a ninth
// particle arrives, children at the float midpoint 0.5*(top+bot).  The
// final shape depends only on the particle set: a cell is internal iff its
// region holds more than 8 particles.  So the device computes, per particle,
// the child digit at every level by descending the same float midpoints,
// radix-sorts the 63-bit digit strings, and cuts the sorted order level by
// level: a node with > 8 particles gets 8 children whose ranges are found
// by binary search on the next digit.  Leaf boxes are the float min/max of
// pos -+ sml (order independent, so bit-exact), empty leaves keep their
// geometric octant, internal boxes are the union of the 8 children.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <vector>

#include "psr_internal.h"

DevState psr_ctx::state() const {
    DevState s;
    const int64_t n = npar;
    s.ps = d_ps;
    s.orig = d_orig;
    s.mass = d_f32 + 0 * n;
    s.rho = d_f32 + 1 * n;
    s.ie = d_f32 + 2 * n;
    s.yeMET = d_f32 + 3 * n;
    s.yGdepHI = d_f32 + 4 * n;
    s.yGdepHeI = d_f32 + 5 * n;
    s.yGdepHeII = d_f32 + 6 * n;
    s.yGrecHII = d_f32 + 7 * n;
    s.yGrecHeII = d_f32 + 8 * n;
    s.yGrecHeIII = d_f32 + 9 * n;
    s.heat = d_f32 + 10 * n;
    s.lH = d_f64 + 0 * n;
    s.lHeI = d_f64 + 1 * n;
    s.lHeII = d_f64 + 2 * n;
    s.lasthit = d_lasthit;
    s.hits = d_hits;
    s.active = active_of(par);
    s.dv = d_dview;
    s.dv_he = dview_he;
    return s;
}

DevPhys psr_dev_phys(const psr_ctx *c) {
    DevPhys p{};
    p.C_H_PER_MASS = c->U.C_H_PER_MASS;
    p.C_HE_PER_MASS = c->U.C_HE_PER_MASS;
    p.C_HI_ENERGY = c->U.C_HI_ENERGY;
    p.C_HEI_ENERGY = c->U.C_HEI_ENERGY;
    p.C_HEII_ENERGY = c->U.C_HEII_ENERGY;
    p.C_HMF = c->U.C_HMF;
    p.C_HEMF = c->U.C_HEMF;
    p.U_MPROTON_OVER_C_BOLTZMANN = c->U.U_MPROTON_OVER_C_BOLTZMANN;
    p.C_BOLTZMANN_OVER_U_MPROTON = c->U.C_BOLTZMANN_OVER_U_MPROTON;
    p.isothermal = c->F.isothermal;
    p.adiabatic = c->F.adiabatic;
    p.on_the_spot = c->F.on_the_spot;
    p.h_only = c->F.h_only;
    p.secondary_ionization = c->F.secondary_ionization;
    p.no_photon = c->F.no_photon;
    const double sf = c->F.comoving ? 1 / (c->F.redshift + 1) : 1.0;
    p.scaling_fac2_inv = 1.0 / (sf * sf);
    const double sf3 = c->F.comoving ? 1 / (c->F.redshift + 1.0) : 1.0;
    p.scaling_fac3_inv = 1.0 / (sf3 * sf3 * sf3);
    p.fake_temperature = c->F.fake_temperature;
    return p;
}

int psr_join(psr_ctx *c) {
    if (!c->upd_pending) return PSR_OK;
    c->upd_pending = false;
    PSR_HIP(c, hipStreamWaitEvent(c->stream, c->ev_upd, 0));
    return PSR_OK;
}

int psr_sync_dev_ctx(psr_ctx *c) {
    if (!c->ctx_dirty && c->d_ctx) return PSR_OK;
    if (int rc = psr_join(c)) return rc;
    if (!c->d_ctx) PSR_HIP(c, hipMalloc(&c->d_ctx, sizeof(DevCtx)));
    DevCtx h;
    h.S = c->state();
    h.T = c->T;
    h.P = psr_dev_phys(c);
    PSR_HIP(c, hipMemcpyAsync(c->d_ctx, &h, sizeof(DevCtx), hipMemcpyHostToDevice, c->stream));
    PSR_HIP(c, hipStreamSynchronize(c->stream));
    c->ctx_dirty = false;
    return PSR_OK;
}

int psr_ensure_scratch(psr_ctx *c, size_t bytes) {
    if (bytes <= c->scratch_bytes) return PSR_OK;
    if (int rc = psr_join(c)) return rc;
    PSR_HIP(c, hipStreamSynchronize(c->stream));  // queued kernels may use the old block
    if (c->d_scratch) PSR_HIP(c, hipFree(c->d_scratch));
    c->d_scratch = nullptr;
    c->scratch_bytes = 0;
    PSR_HIP(c, hipMalloc(&c->d_scratch, bytes));
    c->scratch_bytes = bytes;
    return PSR_OK;
}

template <typename T>
static void dfree(T *&p) {
    if (p) (void)hipFree(p);
    p = nullptr;
}

// ------------------------------------------------------------- context
extern "C" int psr_create(psr_ctx **out, int device) {
    if (!out) return PSR_ERR_ARG;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return PSR_ERR_HIP;
    if (device < 0 || device >= n) return PSR_ERR_ARG;
    if (hipSetDevice(device) != hipSuccess) return PSR_ERR_HIP;
    psr_ctx *c = new psr_ctx();
    c->device = device;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return PSR_ERR_HIP;
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) {
        c->cu_count = prop.multiProcessorCount;
        c->arch = prop.gcnArchName;
    }
    if (hipMalloc(&c->d_stats, sizeof(DevStats) * PSR_NSTAT) != hipSuccess ||
        hipMemset(c->d_stats, 0, sizeof(DevStats) * PSR_NSTAT) != hipSuccess ||
        psr_noalloc_init(device) != PSR_OK) {
        (void)hipFree(c->d_stats);
        (void)hipStreamDestroy(c->stream);
        delete c;
        return PSR_ERR_HIP;
    }
    *out = c;
    return PSR_OK;
}

static void free_tree(psr_ctx *c) {
    dfree(c->d_nodes);
    dfree(c->d_node_level);
    dfree(c->d_field_root);
    dfree(c->d_perm);
    dfree(c->d_ps);
    dfree(c->d_psl);
    dfree(c->d_orig);
    c->have_tree = false;
}
static void free_state(psr_ctx *c) {
    dfree(c->d_f32);
    dfree(c->d_f64);
    dfree(c->d_dview);
    c->dview_valid = false;
    c->dview_he = -1;
    dfree(c->d_lasthit);
    dfree(c->d_hits);
    dfree(c->d_active);
    c->have_state = false;
}

extern "C" void psr_destroy(psr_ctx *c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    (void)psr_join(c);
    (void)hipStreamSynchronize(c->stream);
    if (c->stream2) (void)hipStreamSynchronize(c->stream2);
    if (c->stream3) (void)hipStreamSynchronize(c->stream3);
    if (c->h_ring) (void)hipHostFree(c->h_ring);
    free_tree(c);
    free_state(c);
    psr_mlp_free(c);
    dfree(c->d_tables);
    dfree(c->d_lte);
    dfree(c->d_rays);
    dfree(c->H.d);
    dfree(c->H.b);
    dfree(c->H.slot);
    dfree(c->H.count);
    dfree(c->H.perm);
    dfree(c->H.sslot);
    dfree(c->H.sb);
    dfree(c->H.order);
    dfree(c->d_ray_len);
    dfree(c->d_ray_tm);
    dfree(c->d_bin_rec);
    dfree(c->d_bin_off);
    dfree(c->d_stats);
    dfree(c->d_qlist);
    dfree(c->d_q);
    dfree(c->d_saved);
    dfree(c->d_ctx);
    for (int k = 0; k < 10; k++)
        if (c->ev[k]) (void)hipEventDestroy(c->ev[k]);
    for (int k = 0; k < 2 * psr_ctx::DTUNE_N; k++)
        if (c->tune_ev[k]) (void)hipEventDestroy(c->tune_ev[k]);
    for (int k = 0; k < psr_ctx::TUNE_N; k++)
        if (c->gtune_ev[k]) (void)hipEventDestroy(c->gtune_ev[k]);
    if (c->ev_dep) (void)hipEventDestroy(c->ev_dep);
    if (c->ev_upd) (void)hipEventDestroy(c->ev_upd);
    for (int k = 0; k < 4 * psr_ctx::TTUNE_N; k++)
        if (c->ttune_ev[k]) (void)hipEventDestroy(c->ttune_ev[k]);
    psr_graph_reset(c);
    psr_emit_free(c);
    dfree(c->d_tick_buf);
    c->d_tick = nullptr;
    if (c->d_scratch) (void)hipFree(c->d_scratch);
    if (c->stream2) (void)hipStreamDestroy(c->stream2);
    if (c->stream3) (void)hipStreamDestroy(c->stream3);
    if (c->ev_tr) (void)hipEventDestroy(c->ev_tr);
    if (c->ev_s8) (void)hipEventDestroy(c->ev_s8);
    (void)hipStreamDestroy(c->stream);
    delete c;
}

extern "C" const char *psr_last_error(const psr_ctx *c) { return c ? c->err.c_str() : "null context"; }

extern "C" int psr_sync(psr_ctx *c) {
    if (!c) return PSR_ERR_ARG;
    PSR_HIP(c, hipSetDevice(c->device));
    if (int rc_j = psr_join(c)) return rc_j;
    int rc = psr_resolve(c);
    if (rc) return rc;
    PSR_HIP(c, hipStreamSynchronize(c->stream));
    return PSR_OK;
}

// ------------------------------------------------------------- physics
extern "C" int psr_set_physics(psr_ctx *c, const psr_units *u, const psr_flags *f,
                               const psr_tables *t) {
    if (!c || !u || !f || !t) return PSR_ERR_ARG;
    if (!t->ar || t->ar_nrows < 2 || !t->xs || t->xs_nrows < 2 || !t->kernel || t->kernel_n < 1)
        PSR_FAIL(c, PSR_ERR_ARG, "psr_set_physics: incomplete tables");
    if (f->comoving && f->redshift < 0) PSR_FAIL(c, PSR_ERR_ARG, "negative redshift");
    PSR_HIP(c, hipSetDevice(c->device));
    if (int rc_j = psr_join(c)) return rc_j;
    if (int rc = psr_resolve(c)) return rc;
    dfree(c->d_tables);
    const size_t nar = 16 * (size_t)t->ar_nrows, nxs = 4 * (size_t)t->xs_nrows,
                 nk = (size_t)t->kernel_n;
    const size_t nsi = t->si_n1 > 0 ? (size_t)t->si_n1 + t->si_n2 + 3 * (size_t)t->si_n1 * t->si_n2 : 0;
    PSR_HIP(c, hipMalloc(&c->d_tables, sizeof(double) * (nar + nxs + nk + nsi)));
    double *p = c->d_tables;
    {
        // rate table row-major on the device: one log10 T row (logT and the 15
        // rate columns) is 128 bytes, a single cache line per interpolation end
        std::vector<double> rows(nar);
        for (int32_t i = 0; i < t->ar_nrows; i++)
            for (int k = 0; k < 16; k++) rows[(size_t)i * 16 + k] = t->ar[(size_t)k * t->ar_nrows + i];
        PSR_HIP(c, hipMemcpy(p, rows.data(), sizeof(double) * nar, hipMemcpyHostToDevice));
    }
    PSR_HIP(c, hipMemcpy(p + nar, t->xs, sizeof(double) * nxs, hipMemcpyHostToDevice));
    PSR_HIP(c, hipMemcpy(p + nar + nxs, t->kernel, sizeof(double) * nk, hipMemcpyHostToDevice));
    DevTables T{};
    T.ar = p;
    T.ar_nrows = t->ar_nrows;
    T.ar_min = t->ar_min;
    T.ar_step_inv = t->ar_step_inv;
    T.xs = p + nar;
    T.xs_nrows = t->xs_nrows;
    T.xs_min = t->xs_min;
    T.xs_step_inv = t->xs_step_inv;
    for (int k = 0; k < 3; k++) T.xs_cut[k] = t->xs_cut[k];
    T.kernel = p + nar + nxs;
    T.kernel_n = t->kernel_n;
    if (nsi) {
        double *q = p + nar + nxs + nk;
        const size_t n12 = (size_t)t->si_n1 * t->si_n2;
        PSR_HIP(c, hipMemcpy(q, t->si_x, sizeof(double) * t->si_n1, hipMemcpyHostToDevice));
        PSR_HIP(c, hipMemcpy(q + t->si_n1, t->si_y, sizeof(double) * t->si_n2, hipMemcpyHostToDevice));
        double *d = q + t->si_n1 + t->si_n2;
        {
            // the three tables interleaved: one lookup reads all three from the
            // same cache lines
            std::vector<double> tab3(3 * n12);
            for (size_t k = 0; k < n12; k++) {
                tab3[3 * k] = t->si_phi_hi[k];
                tab3[3 * k + 1] = t->si_phi_hei[k];
                tab3[3 * k + 2] = t->si_eh[k];
            }
            PSR_HIP(c, hipMemcpy(d, tab3.data(), sizeof(double) * 3 * n12, hipMemcpyHostToDevice));
        }
        T.si_n1 = t->si_n1;
        T.si_n2 = t->si_n2;
        T.si_min1 = t->si_min1;
        T.si_min2 = t->si_min2;
        T.si_step1_inv = t->si_step1_inv;
        T.si_step2_inv = t->si_step2_inv;
        T.si_x = q;
        T.si_y = q + t->si_n1;
        T.si_tab3 = d;
    }
    c->T = T;
    // LTE recombination spectra for the recombination rays (src/lte.c:28-72)
    {
        const size_t ne = 3 * (size_t)PSR_LTE_NFREQ, nc = ne * PSR_LTE_NLOGT, nb = PSR_LTE_NLOGT - 1;
        std::vector<double> lte(ne + nc + nb);
        if (psr_lte_tables(t, u, lte.data(), lte.data() + ne, lte.data() + ne + nc))
            PSR_FAIL(c, PSR_ERR_ARG, "psr_set_physics: LTE spectra");
        if (!c->d_lte) PSR_HIP(c, hipMalloc(&c->d_lte, sizeof(double) * lte.size()));
        PSR_HIP(c, hipMemcpy(c->d_lte, lte.data(), sizeof(double) * lte.size(), hipMemcpyHostToDevice));
    }
    c->U = *u;
    c->F = *f;
    c->have_physics = true;
    c->ctx_dirty = true;
    psr_tune_reset(c);
    psr_graph_reset(c);
    return PSR_OK;
}

// ------------------------------------------------------------ tree build
// child digit of every level: bit d set <=> pos[d] >= float midpoint, the
// rule of split()/inside() (src/raytrace.c:280-287, 306-316)
__global__ void k_keys(const float *__restrict__ pos, int64_t n, const int64_t *__restrict__ fo,
                       int32_t nfield, const float *__restrict__ boxf,
                       unsigned long long *__restrict__ keys, uint32_t *__restrict__ idx,
                       uint16_t *__restrict__ fid) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    int lo = 0, hi = nfield - 1;
    while (lo < hi) {  // field whose range holds i
        const int mid = (lo + hi + 1) >> 1;
        if (fo[mid] <= i) lo = mid; else hi = mid - 1;
    }
    const float p[3] = {pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]};
    float bot[3] = {0.f, 0.f, 0.f};
    const float B = boxf[lo];
    float top[3] = {B, B, B};
    unsigned long long key = 0;
    bool inside = true;
    for (int d = 0; d < 3; d++)
        if (p[d] >= top[d] || p[d] < bot[d]) inside = false;
    if (inside) {
        for (int l = 0; l < PSR_MAXDEPTH; l++) {
            unsigned dig = 0;
            for (int d = 0; d < 3; d++) {
                const float ctr = (float)(0.5 * (double)(top[d] + bot[d]));
                if (p[d] >= ctr) {
                    dig |= 1u << d;
                    bot[d] = ctr;
                } else {
                    top[d] = ctr;
                }
            }
            key = (key << 3) | dig;
        }
    } else {
        key = ~0ull;  // outside the root: skipped (src/raytrace.c:70-75)
    }
    keys[i] = key;
    idx[i] = (uint32_t)i;
    fid[i] = (uint16_t)lo;
}

__global__ void k_gather_fid(const uint32_t *__restrict__ idx, const uint16_t *__restrict__ fid_in,
                             uint16_t *__restrict__ fid_out, int64_t n) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n) fid_out[i] = fid_in[idx[i]];
}

__global__ void k_gather_keys(const uint32_t *__restrict__ idx,
                              const unsigned long long *__restrict__ kin,
                              unsigned long long *__restrict__ kout, int64_t n) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n) kout[i] = kin[idx[i]];
}

// tree-order particle records: ps = (x, y, z, sml), orig = index | invalid.
// An invalid particle's sml is stored negated (its sign bit is the flag, so
// the traversal loads one 16-byte record per particle); w * w and |w| are
// unchanged
__global__ void k_particles(const uint32_t *__restrict__ idx, const float *__restrict__ pos,
                            const float *__restrict__ sml, const int8_t *__restrict__ flag,
                            float4 *__restrict__ ps, uint32_t *__restrict__ orig, int64_t n) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t j = idx[i];
    const bool invalid = flag && (flag[j] & 1);
    const float h = fabsf(sml[j]);
    ps[i] = make_float4(pos[3 * (size_t)j], pos[3 * (size_t)j + 1], pos[3 * (size_t)j + 2], invalid ? -h : h);
    orig[i] = j | (invalid ? PSR_INVALID_BIT : 0u);
}

// roots: one per field, range = valid particles of the field
__global__ void k_roots(const unsigned long long *__restrict__ keys, const int64_t *__restrict__ fo,
                        int32_t nfield, const float *__restrict__ boxf, int32_t *__restrict__ nd_start,
                        int32_t *__restrict__ nd_count, float *__restrict__ nd_gb,
                        float *__restrict__ nd_gt, int32_t *__restrict__ nd_level,
                        int32_t *__restrict__ nd_child, int32_t *__restrict__ root) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nfield) return;
    int64_t lo = fo[f], hi = fo[f + 1];
    while (lo < hi) {  // first invalid key of the segment
        const int64_t mid = (lo + hi) >> 1;
        if (keys[mid] == ~0ull) hi = mid; else lo = mid + 1;
    }
    nd_start[f] = (int32_t)fo[f];
    nd_count[f] = (int32_t)(lo - fo[f]);
    for (int d = 0; d < 3; d++) {
        nd_gb[3 * f + d] = 0.f;
        nd_gt[3 * f + d] = boxf[f];
    }
    nd_level[f] = 0;
    nd_child[f] = -1;
    root[f] = f;
}

__global__ void k_flag_internal(const int32_t *__restrict__ nd_count, int32_t ls, int32_t n,
                                int32_t *__restrict__ flag) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < n) flag[j] = nd_count[ls + j] > 8 ? 1 : 0;
}

// children of the internal nodes of one level
__global__ void k_children(const unsigned long long *__restrict__ keys, int32_t ls, int32_t n,
                           int32_t next, const int32_t *__restrict__ flag,
                           const int32_t *__restrict__ scan, int32_t level, int32_t *nd_start,
                           int32_t *nd_count, float *nd_gb, float *nd_gt, int32_t *nd_level,
                           int32_t *nd_child) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = t >> 3, c = t & 7;
    if (j >= n || !flag[j]) return;
    const int node = ls + j;
    const int base = next + 8 * scan[j];
    const int child = base + c;
    if (c == 0) nd_child[node] = base;
    const int s = nd_start[node], cnt = nd_count[node];
    const int shift = 3 * (PSR_MAXDEPTH - (level + 1));
    auto lower = [&](unsigned v) {  // first position whose digit >= v
        int lo = s, hi = s + cnt;
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (((keys[mid] >> shift) & 7u) >= v) hi = mid; else lo = mid + 1;
        }
        return lo;
    };
    const int a = lower((unsigned)c), b = (c == 7) ? s + cnt : lower((unsigned)c + 1);
    nd_start[child] = a;
    nd_count[child] = b - a;
    for (int d = 0; d < 3; d++) {
        const float pb = nd_gb[3 * node + d], pt = nd_gt[3 * node + d];
        const float ctr = (float)(0.5 * (double)(pt + pb));
        nd_gb[3 * child + d] = ((1 << d) & c) ? ctr : pb;
        nd_gt[3 * child + d] = ((1 << d) & c) ? pt : ctr;
    }
    nd_level[child] = level + 1;
    nd_child[child] = -1;
}

// leaf boxes (src/raytrace.c:117-127): union of [pos - sml, pos + sml] in
// float; an empty leaf keeps its octant box
__global__ void k_leaf_aabb(const float4 *__restrict__ ps, int32_t nn,
                            const int32_t *__restrict__ nd_start, const int32_t *__restrict__ nd_count,
                            const float *__restrict__ nd_gb, const float *__restrict__ nd_gt,
                            const int32_t *__restrict__ nd_child, Node *__restrict__ nodes) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nn || nd_child[j] != -1) return;
    Node o;
    const int s = nd_start[j], cnt = nd_count[j];
    if (cnt == 0) {
        for (int d = 0; d < 3; d++) {
            o.bot[d] = nd_gb[3 * j + d];
            o.top[d] = nd_gt[3 * j + d];
        }
    } else {
        for (int k = 0; k < cnt; k++) {
            const float4 p = ps[s + k];
            const float q[3] = {p.x, p.y, p.z};
            const float h = fabsf(p.w);
            for (int d = 0; d < 3; d++) {
                const float lo = q[d] - h, hi = q[d] + h;
                if (k == 0 || lo < o.bot[d]) o.bot[d] = lo;
                if (k == 0 || hi > o.top[d]) o.top[d] = hi;
            }
        }
    }
    o.a = s;
    o.b = cnt;
    nodes[j] = o;
}

// internal boxes (src/raytrace.c:128-141): union of the 8 children
__global__ void k_internal_aabb(int32_t ls, int32_t n, const int32_t *__restrict__ nd_child,
                                Node *__restrict__ nodes) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const int node = ls + j;
    const int fc = nd_child[node];
    if (fc < 0) return;
    Node o;
    for (int i = 0; i < 8; i++) {
        const Node ch = nodes[fc + i];
        for (int d = 0; d < 3; d++) {
            if (i == 0 || ch.bot[d] < o.bot[d]) o.bot[d] = ch.bot[d];
            if (i == 0 || ch.top[d] > o.top[d]) o.top[d] = ch.top[d];
        }
    }
    o.a = fc;
    o.b = -1;
    nodes[node] = o;
}

// the trace's particle lines: slot s of leaf j is its particle s, or an
// invalid record (sml -0: sign bit set) past its count
__global__ void k_leaf_lines(const float4 *__restrict__ ps, const Node *__restrict__ nodes, int32_t nn,
                             float4 *__restrict__ psl) {
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    const int64_t j = t >> 3;
    const int s = (int)(t & 7);
    if (j >= nn) return;
    const Node nd = nodes[j];
    if (nd.b <= 0) return;
    psl[t] = s < nd.b ? ps[nd.a + s] : make_float4(0.0f, 0.0f, 0.0f, -0.0f);
}

static inline unsigned nblk(int64_t n, int t = 256) { return (unsigned)((n + t - 1) / t); }

extern "C" int psr_tree_build(psr_ctx *c, const float *pos, const float *sml, const int8_t *flag,
                              int64_t npar, int32_t nfield, const int64_t *field_offsets,
                              const double *boxsizes) {
    if (!c || npar < 0 || nfield < 1 || nfield > 65535 || !field_offsets || !boxsizes)
        return PSR_ERR_ARG;
    if (npar > 0 && (!pos || !sml)) return PSR_ERR_ARG;
    if (npar >= (int64_t)INT32_MAX) PSR_FAIL(c, PSR_ERR_ARG, "npar must be < 2^31");
    if (field_offsets[0] != 0 || field_offsets[nfield] != npar)
        PSR_FAIL(c, PSR_ERR_ARG, "field_offsets must run from 0 to npar");
    for (int f = 0; f < nfield; f++) {
        if (field_offsets[f + 1] < field_offsets[f]) PSR_FAIL(c, PSR_ERR_ARG, "field_offsets decrease");
        if (!(boxsizes[f] > 0)) PSR_FAIL(c, PSR_ERR_ARG, "boxsize must be > 0");
    }
    PSR_HIP(c, hipSetDevice(c->device));
    if (int rc_j = psr_join(c)) return rc_j;
    if (int rc = psr_resolve(c)) return rc;
    hipStream_t st = c->stream;
    free_tree(c);
    free_state(c);
    c->hits_nray = 0;  // the hit lists name slots of the old tree
    c->npar = npar;
    c->nfield = nfield;
    c->boxsizes.assign(boxsizes, boxsizes + nfield);
    const int64_t n = npar;
    const int64_t nal = n > 0 ? n : 1;

    // inputs
    float *d_pos = nullptr, *d_sml = nullptr, *d_boxf = nullptr;
    int8_t *d_flag = nullptr;
    int64_t *d_fo = nullptr;
    unsigned long long *d_k0 = nullptr, *d_k1 = nullptr;
    uint32_t *d_i0 = nullptr, *d_i1 = nullptr;
    uint16_t *d_f0 = nullptr, *d_f1 = nullptr;
    std::vector<float> boxf(nfield);
    for (int f = 0; f < nfield; f++) boxf[f] = (float)boxsizes[f];
    auto cleanup = [&]() {
        dfree(d_pos); dfree(d_sml); dfree(d_boxf); dfree(d_flag); dfree(d_fo);
        dfree(d_k0); dfree(d_k1); dfree(d_i0); dfree(d_i1); dfree(d_f0); dfree(d_f1);
    };
#define TB_HIP(call)                                                             \
    do {                                                                         \
        hipError_t e_ = (call);                                                  \
        if (e_ != hipSuccess) {                                                  \
            c->err = std::string(#call) + ": " + hipGetErrorString(e_);          \
            cleanup();                                                           \
            return PSR_ERR_HIP;                                                  \
        }                                                                        \
    } while (0)
    TB_HIP(hipMalloc(&d_pos, sizeof(float) * 3 * nal));
    TB_HIP(hipMalloc(&d_sml, sizeof(float) * nal));
    TB_HIP(hipMalloc(&d_boxf, sizeof(float) * nfield));
    TB_HIP(hipMalloc(&d_fo, sizeof(int64_t) * (nfield + 1)));
    if (flag) TB_HIP(hipMalloc(&d_flag, nal));
    TB_HIP(hipMalloc(&d_k0, sizeof(unsigned long long) * nal));
    TB_HIP(hipMalloc(&d_k1, sizeof(unsigned long long) * nal));
    TB_HIP(hipMalloc(&d_i0, sizeof(uint32_t) * nal));
    TB_HIP(hipMalloc(&d_i1, sizeof(uint32_t) * nal));
    TB_HIP(hipMalloc(&d_f0, sizeof(uint16_t) * nal));
    TB_HIP(hipMalloc(&d_f1, sizeof(uint16_t) * nal));
    if (n) {
        TB_HIP(hipMemcpyAsync(d_pos, pos, sizeof(float) * 3 * n, hipMemcpyHostToDevice, st));
        TB_HIP(hipMemcpyAsync(d_sml, sml, sizeof(float) * n, hipMemcpyHostToDevice, st));
        if (flag) TB_HIP(hipMemcpyAsync(d_flag, flag, n, hipMemcpyHostToDevice, st));
    }
    TB_HIP(hipMemcpyAsync(d_boxf, boxf.data(), sizeof(float) * nfield, hipMemcpyHostToDevice, st));
    TB_HIP(hipMemcpyAsync(d_fo, field_offsets, sizeof(int64_t) * (nfield + 1), hipMemcpyHostToDevice, st));

    // keys, then sort by key and (stably) by field
    if (n) {
        k_keys<<<nblk(n), 256, 0, st>>>(d_pos, n, d_fo, nfield, d_boxf, d_k0, d_i0, d_f0);
        TB_HIP(hipGetLastError());
        // the keys are octree paths of 3 bits per level below the root
        // (PSR_MAXDEPTH levels, 63 bits); particles outside the root carry
        // all 64 bits set and sort last
        const int kbits = 64;
        const size_t tb = std::max(psr_sort_scratch_u64(n), psr_sort_scratch_u16(n));
        if (psr_ensure_scratch(c, tb)) { cleanup(); return PSR_ERR_HIP; }
        TB_HIP(psr_sort_pairs_u64(d_k0, d_i0, d_k1, d_i1, n, 0, kbits, c->d_scratch, st));
        if (nfield > 1) {
            k_gather_fid<<<nblk(n), 256, 0, st>>>(d_i1, d_f0, d_f1, n);
            int bits = 1;
            while ((1 << bits) < nfield) bits++;
            TB_HIP(psr_sort_pairs_u16(d_f1, d_i1, d_f0, d_i0, n, 0, bits, c->d_scratch, st));
            // d_i0 holds the final order; regather keys from the unsorted copy
            k_gather_keys<<<nblk(n), 256, 0, st>>>(d_i0, d_k0, d_k1, n);
            TB_HIP(hipMemcpyAsync(d_i1, d_i0, sizeof(uint32_t) * n, hipMemcpyDeviceToDevice, st));
        }
    }
    // d_k1: sorted keys, d_i1: tree slot -> caller index
    TB_HIP(hipMalloc(&c->d_ps, sizeof(float4) * nal));
    TB_HIP(hipMalloc(&c->d_orig, sizeof(uint32_t) * nal));
    if (n) {
        k_particles<<<nblk(n), 256, 0, st>>>(d_i1, d_pos, d_sml, d_flag, c->d_ps, c->d_orig, n);
        TB_HIP(hipGetLastError());
    }

    // level-by-level node construction
    int64_t cap = std::max<int64_t>(1024, n / 2 + 8 * nfield + 64);
    int32_t *nd_start = nullptr, *nd_count = nullptr, *nd_level = nullptr, *nd_child = nullptr;
    float *nd_gb = nullptr, *nd_gt = nullptr;
    int32_t *d_flagi = nullptr, *d_scan = nullptr;
    auto cleanup2 = [&]() {
        dfree(nd_start); dfree(nd_count); dfree(nd_level); dfree(nd_child);
        dfree(nd_gb); dfree(nd_gt); dfree(d_flagi); dfree(d_scan);
    };
#undef TB_HIP
#define TB_HIP(call)                                                             \
    do {                                                                         \
        hipError_t e_ = (call);                                                  \
        if (e_ != hipSuccess) {                                                  \
            c->err = std::string(#call) + ": " + hipGetErrorString(e_);          \
            cleanup();                                                           \
            cleanup2();                                                          \
            return PSR_ERR_HIP;                                                  \
        }                                                                        \
    } while (0)
    auto alloc_nodes = [&](int64_t newcap, int64_t used) -> hipError_t {
        int32_t *a = nullptr, *b = nullptr, *l = nullptr, *ch = nullptr;
        float *gb = nullptr, *gt = nullptr;
        hipError_t e;
        if ((e = hipMalloc(&a, sizeof(int32_t) * newcap)) != hipSuccess) return e;
        if ((e = hipMalloc(&b, sizeof(int32_t) * newcap)) != hipSuccess) return e;
        if ((e = hipMalloc(&l, sizeof(int32_t) * newcap)) != hipSuccess) return e;
        if ((e = hipMalloc(&ch, sizeof(int32_t) * newcap)) != hipSuccess) return e;
        if ((e = hipMalloc(&gb, sizeof(float) * 3 * newcap)) != hipSuccess) return e;
        if ((e = hipMalloc(&gt, sizeof(float) * 3 * newcap)) != hipSuccess) return e;
        if (used) {
            (void)hipMemcpyAsync(a, nd_start, sizeof(int32_t) * used, hipMemcpyDeviceToDevice, st);
            (void)hipMemcpyAsync(b, nd_count, sizeof(int32_t) * used, hipMemcpyDeviceToDevice, st);
            (void)hipMemcpyAsync(l, nd_level, sizeof(int32_t) * used, hipMemcpyDeviceToDevice, st);
            (void)hipMemcpyAsync(ch, nd_child, sizeof(int32_t) * used, hipMemcpyDeviceToDevice, st);
            (void)hipMemcpyAsync(gb, nd_gb, sizeof(float) * 3 * used, hipMemcpyDeviceToDevice, st);
            (void)hipMemcpyAsync(gt, nd_gt, sizeof(float) * 3 * used, hipMemcpyDeviceToDevice, st);
            (void)hipStreamSynchronize(st);
        }
        dfree(nd_start); dfree(nd_count); dfree(nd_level); dfree(nd_child); dfree(nd_gb); dfree(nd_gt);
        nd_start = a; nd_count = b; nd_level = l; nd_child = ch; nd_gb = gb; nd_gt = gt;
        cap = newcap;
        return hipSuccess;
    };
    TB_HIP(alloc_nodes(cap, 0));
    TB_HIP(hipMalloc(&c->d_field_root, sizeof(int32_t) * nfield));
    k_roots<<<nblk(nfield), 256, 0, st>>>(d_k1, d_fo, nfield, d_boxf, nd_start, nd_count, nd_gb, nd_gt,
                                         nd_level, nd_child, c->d_field_root);
    TB_HIP(hipGetLastError());
    std::vector<int32_t> level_start;
    level_start.push_back(0);
    int32_t ls = 0, le = nfield;
    int level = 0;
    int64_t flag_cap = 0;
    for (;;) {
        const int32_t nl = le - ls;
        if (nl > flag_cap) {
            dfree(d_flagi);
            dfree(d_scan);
            flag_cap = std::max<int64_t>(nl, 2 * flag_cap);
            TB_HIP(hipMalloc(&d_flagi, sizeof(int32_t) * flag_cap));
            TB_HIP(hipMalloc(&d_scan, sizeof(int32_t) * flag_cap));
        }
        k_flag_internal<<<nblk(nl), 256, 0, st>>>(nd_count, ls, nl, d_flagi);
        if (psr_ensure_scratch(c, psr_scan_scratch_i32(nl))) { cleanup(); cleanup2(); return PSR_ERR_HIP; }
        TB_HIP(psr_exclusive_sum_i32(d_flagi, d_scan, nl, c->d_scratch, st));
        int32_t last_scan = 0, last_flag = 0;
        TB_HIP(hipMemcpyAsync(&last_scan, d_scan + nl - 1, sizeof(int32_t), hipMemcpyDeviceToHost, st));
        TB_HIP(hipMemcpyAsync(&last_flag, d_flagi + nl - 1, sizeof(int32_t), hipMemcpyDeviceToHost, st));
        TB_HIP(hipStreamSynchronize(st));
        const int64_t ninternal = (int64_t)last_scan + last_flag;
        if (ninternal == 0) break;
        if (level + 1 > PSR_MAXDEPTH) {
            cleanup();
            cleanup2();
            PSR_FAIL(c, PSR_ERR_TREE, "octree deeper than 21 levels (more than 8 particles within "
                                      "one depth-21 cell)");
        }
        const int64_t need = (int64_t)le + 8 * ninternal;
        if (need >= INT32_MAX) { cleanup(); cleanup2(); PSR_FAIL(c, PSR_ERR_TREE, "too many nodes"); }
        if (need > cap) TB_HIP(alloc_nodes(std::max<int64_t>(need, 2 * cap), le));
        k_children<<<nblk(8 * (int64_t)nl), 256, 0, st>>>(d_k1, ls, nl, le, d_flagi, d_scan, level,
                                                          nd_start, nd_count, nd_gb, nd_gt, nd_level,
                                                          nd_child);
        TB_HIP(hipGetLastError());
        ls = le;
        le = (int32_t)need;
        level++;
        level_start.push_back(ls);
    }
    level_start.push_back(le);
    const int32_t nn = le;
    c->nnode = nn;
    c->depth = level;
    TB_HIP(hipMalloc(&c->d_nodes, sizeof(Node) * nn));
    TB_HIP(hipMalloc(&c->d_node_level, sizeof(int32_t) * nn));
    TB_HIP(hipMemcpyAsync(c->d_node_level, nd_level, sizeof(int32_t) * nn, hipMemcpyDeviceToDevice, st));
    k_leaf_aabb<<<nblk(nn), 256, 0, st>>>(c->d_ps, nn, nd_start, nd_count, nd_gb, nd_gt, nd_child,
                                         c->d_nodes);
    TB_HIP(hipGetLastError());
    for (int l = level - 1; l >= 0; l--) {
        const int32_t a = level_start[l], b = level_start[l + 1];
        k_internal_aabb<<<nblk(b - a), 256, 0, st>>>(a, b - a, nd_child, c->d_nodes);
        TB_HIP(hipGetLastError());
    }
    TB_HIP(hipMalloc(&c->d_psl, sizeof(float4) * 8 * (size_t)nn));
    k_leaf_lines<<<nblk(8 * (int64_t)nn), 256, 0, st>>>(c->d_ps, c->d_nodes, nn, c->d_psl);
    TB_HIP(hipGetLastError());
    // skipped particles: invalid keys
    std::vector<int32_t> rootcnt(nfield);
    TB_HIP(hipMemcpyAsync(rootcnt.data(), nd_count, sizeof(int32_t) * nfield, hipMemcpyDeviceToHost, st));
    TB_HIP(hipStreamSynchronize(st));
    int64_t valid = 0;
    for (int f = 0; f < nfield; f++) valid += rootcnt[f];
    c->skipped = n - valid;
    // keep the permutation, drop temporaries
    c->d_perm = d_i1;
    d_i1 = nullptr;
    cleanup();
    cleanup2();
#undef TB_HIP
    c->have_tree = true;
    c->ctx_dirty = true;
    psr_tune_reset(c);
    psr_graph_reset(c);
    return PSR_OK;
}

extern "C" int psr_tree_info(psr_ctx *c, int64_t *nnodes, int32_t *depth, int64_t *skipped) {
    if (!c) return PSR_ERR_ARG;
    if (!c->have_tree) PSR_FAIL(c, PSR_ERR_STATE, "no tree");
    if (nnodes) *nnodes = c->nnode;
    if (depth) *depth = c->depth;
    if (skipped) *skipped = c->skipped;
    return PSR_OK;
}

extern "C" int psr_tree_export(psr_ctx *c, int64_t cap, float *bot, float *top, int32_t *first_child,
                               int32_t *npar, int32_t *level) {
    if (!c) return PSR_ERR_ARG;
    if (!c->have_tree) PSR_FAIL(c, PSR_ERR_STATE, "no tree");
    if (cap < c->nnode) PSR_FAIL(c, PSR_ERR_CAPACITY, "export buffer smaller than node count");
    std::vector<Node> h(c->nnode);
    PSR_HIP(c, hipMemcpy(h.data(), c->d_nodes, sizeof(Node) * c->nnode, hipMemcpyDeviceToHost));
    if (level) PSR_HIP(c, hipMemcpy(level, c->d_node_level, sizeof(int32_t) * c->nnode, hipMemcpyDeviceToHost));
    for (int64_t i = 0; i < c->nnode; i++) {
        for (int d = 0; d < 3; d++) {
            if (bot) bot[3 * i + d] = h[i].bot[d];
            if (top) top[3 * i + d] = h[i].top[d];
        }
        if (first_child) first_child[i] = h[i].b < 0 ? h[i].a : -1;
        if (npar) npar[i] = h[i].b < 0 ? -1 : h[i].b;
    }
    return PSR_OK;
}

// ------------------------------------------------------------ state
template <typename T>
__global__ void k_gather(const uint32_t *__restrict__ perm, const T *__restrict__ src,
                         T *__restrict__ dst, int64_t n) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[perm[i]];
}
template <typename T>
__global__ void k_scatter(const uint32_t *__restrict__ perm, const T *__restrict__ src,
                          T *__restrict__ dst, int64_t n) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n) dst[perm[i]] = src[i];
}

template <typename T>
static int upload_field(psr_ctx *c, const T *host, T *dst, void *tmp) {
    const int64_t n = c->npar;
    if (!n) return PSR_OK;
    if (!host) PSR_FAIL(c, PSR_ERR_ARG, "psr_state_upload: missing field");
    PSR_HIP(c, hipMemcpyAsync(tmp, host, sizeof(T) * n, hipMemcpyHostToDevice, c->stream));
    k_gather<T><<<nblk(n), 256, 0, c->stream>>>(c->d_perm, (const T *)tmp, dst, n);
    PSR_HIP(c, hipGetLastError());
    return PSR_OK;
}
template <typename T>
static int download_field(psr_ctx *c, T *host, const T *src, void *tmp) {
    const int64_t n = c->npar;
    if (!n || !host) return PSR_OK;
    k_scatter<T><<<nblk(n), 256, 0, c->stream>>>(c->d_perm, src, (T *)tmp, n);
    PSR_HIP(c, hipGetLastError());
    PSR_HIP(c, hipMemcpyAsync(host, tmp, sizeof(T) * n, hipMemcpyDeviceToHost, c->stream));
    PSR_HIP(c, hipStreamSynchronize(c->stream));
    return PSR_OK;
}

extern "C" int psr_state_upload(psr_ctx *c, const psr_state *s) {
    if (!c || !s) return PSR_ERR_ARG;
    if (!c->have_tree) PSR_FAIL(c, PSR_ERR_STATE, "psr_state_upload before psr_tree_build");
    if (s->npar != c->npar) PSR_FAIL(c, PSR_ERR_ARG, "psr_state_upload: npar differs from the tree");
    PSR_HIP(c, hipSetDevice(c->device));
    if (int rc_j = psr_join(c)) return rc_j;
    if (int rc = psr_resolve(c)) return rc;
    const int64_t n = c->npar, nal = n > 0 ? n : 1;
    if (!c->have_state) {
        PSR_HIP(c, hipMalloc(&c->d_f32, sizeof(float) * 11 * nal));
        PSR_HIP(c, hipMalloc(&c->d_f64, sizeof(double) * 3 * nal));
        PSR_HIP(c, hipMalloc(&c->d_dview, sizeof(DepViewHe) * nal));
        PSR_HIP(c, hipMalloc(&c->d_lasthit, sizeof(int64_t) * nal));
        PSR_HIP(c, hipMalloc(&c->d_hits, sizeof(int32_t) * nal));
        PSR_HIP(c, hipMalloc(&c->d_active, 2 * nal));
        PSR_HIP(c, hipMemsetAsync(c->d_active, 0, 2 * nal, c->stream));
    }
    if (psr_ensure_scratch(c, sizeof(double) * nal)) return PSR_ERR_HIP;
    void *tmp = c->d_scratch;
    DevState d = c->state();
    int rc;
#define UP(field, dst)                                                   \
    if ((rc = upload_field(c, (decltype(+*(dst)) *)(s->field), dst, tmp))) return rc;
    UP(mass, d.mass);
    UP(rho, d.rho);
    UP(ie, d.ie);
    UP(yeMET, d.yeMET);
    UP(yGdepHI, d.yGdepHI);
    UP(yGdepHeI, d.yGdepHeI);
    UP(yGdepHeII, d.yGdepHeII);
    UP(yGrecHII, d.yGrecHII);
    UP(yGrecHeII, d.yGrecHeII);
    UP(yGrecHeIII, d.yGrecHeIII);
    UP(heat, d.heat);
    UP(lambdaH, d.lH);
    UP(lambdaHeI, d.lHeI);
    UP(lambdaHeII, d.lHeII);
    UP(lasthit, d.lasthit);
    UP(hits, d.hits);
#undef UP
    PSR_HIP(c, hipStreamSynchronize(c->stream));
    c->have_state = true;
    c->dview_valid = false;   // rebuilt before the next deposit
    c->ctx_dirty = true;
    psr_graph_reset(c);
    return PSR_OK;
}

extern "C" int psr_state_download(psr_ctx *c, psr_state *s) {
    if (!c || !s) return PSR_ERR_ARG;
    if (!c->have_state) PSR_FAIL(c, PSR_ERR_STATE, "no particle state");
    if (s->npar != c->npar) PSR_FAIL(c, PSR_ERR_ARG, "psr_state_download: npar differs");
    PSR_HIP(c, hipSetDevice(c->device));
    if (int rc_j = psr_join(c)) return rc_j;
    if (int rc = psr_resolve(c)) return rc;
    const int64_t nal = c->npar > 0 ? c->npar : 1;
    if (psr_ensure_scratch(c, sizeof(double) * 3 * nal)) return PSR_ERR_HIP;
    void *tmp = c->d_scratch;
    DevState d = c->state();
    int rc;
#define DN(field, src) \
    if ((rc = download_field(c, s->field, src, tmp))) return rc;
    DN(mass, d.mass);
    DN(rho, d.rho);
    DN(ie, d.ie);
    DN(yeMET, d.yeMET);
    DN(yGdepHI, d.yGdepHI);
    DN(yGdepHeI, d.yGdepHeI);
    DN(yGdepHeII, d.yGdepHeII);
    DN(yGrecHII, d.yGrecHII);
    DN(yGrecHeII, d.yGrecHeII);
    DN(yGrecHeIII, d.yGrecHeIII);
    DN(heat, d.heat);
    DN(lambdaH, d.lH);
    DN(lambdaHeI, d.lHeI);
    DN(lambdaHeII, d.lHeII);
    DN(lasthit, d.lasthit);
    DN(hits, d.hits);
#undef DN
    return PSR_OK;
}

// the blocks of an output snapshot the ticks change, as
// psystem_write_output derives them (src/psystem.c:644-667): the fractions
// and ye (lambda_to_ye) as floats, yGrecHII as gammaHI, ie, lasthit, hits,
// each written in the caller's order into one of eight columns of out
// (4-byte columns of n entries, lasthit an 8-byte column)
__global__ void k_output(DevState S, DevPhys P, const uint32_t *__restrict__ perm, int64_t n, float *__restrict__ f32,
                         int64_t *__restrict__ lasthit, int32_t *__restrict__ hits) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int64_t j = perm[i];
    const double lH = S.lH[i], lHeI = S.lHeI[i], lHeII = S.lHeII[i];
    f32[0 * n + j] = (float)lH;
    f32[1 * n + j] = (float)lHeI;
    f32[2 * n + j] = (float)lHeII;
    f32[3 * n + j] = (float)lambda_ye(P, lH, lHeI, lHeII);
    f32[4 * n + j] = S.yGrecHII[i];
    f32[5 * n + j] = S.ie[i];
    lasthit[j] = S.lasthit[i];
    hits[j] = S.hits[i];
}

extern "C" int psr_output_gather(psr_ctx *c, float *xHI, float *xHeI, float *xHeII, float *ye, float *gammaHI,
                                 float *ie, int64_t *lasthit, int32_t *hits) {
    if (!c) return PSR_ERR_ARG;
    if (!c->have_state) PSR_FAIL(c, PSR_ERR_STATE, "no particle state");
    PSR_HIP(c, hipSetDevice(c->device));
    if (int rc_j = psr_join(c)) return rc_j;
    if (int rc = psr_resolve(c)) return rc;
    const int64_t n = c->npar;
    if (n == 0) return PSR_OK;
    // six float columns, then lasthit, then hits
    if (psr_ensure_scratch(c, (sizeof(float) * 6 + sizeof(int64_t) + sizeof(int32_t)) * n + 64)) return PSR_ERR_HIP;
    float *f32 = (float *)c->d_scratch;
    int64_t *lh = (int64_t *)(((uintptr_t)(f32 + 6 * n) + 7) & ~(uintptr_t)7);
    int32_t *ht = (int32_t *)(lh + n);
    k_output<<<nblk(n), 256, 0, c->stream>>>(c->state(), psr_dev_phys(c), c->d_perm, n, f32, lh, ht);
    PSR_HIP(c, hipGetLastError());
    float *dst[6] = {xHI, xHeI, xHeII, ye, gammaHI, ie};
    for (int k = 0; k < 6; k++)
        if (dst[k]) PSR_HIP(c, hipMemcpyAsync(dst[k], f32 + k * n, sizeof(float) * n, hipMemcpyDeviceToHost, c->stream));
    if (lasthit) PSR_HIP(c, hipMemcpyAsync(lasthit, lh, sizeof(int64_t) * n, hipMemcpyDeviceToHost, c->stream));
    if (hits) PSR_HIP(c, hipMemcpyAsync(hits, ht, sizeof(int32_t) * n, hipMemcpyDeviceToHost, c->stream));
    PSR_HIP(c, hipStreamSynchronize(c->stream));
    return PSR_OK;
}

// ------------------------------------------------------------ rays
extern "C" int psr_rays_upload(psr_ctx *c, const psr_ray *rays, int64_t nray) {
    if (!c || nray < 0 || (nray > 0 && !rays)) return PSR_ERR_ARG;
    PSR_HIP(c, hipSetDevice(c->device));
    if (int rc_j = psr_join(c)) return rc_j;
    // a replay of a pending tick needs the pool it was launched with
    if (int rc = psr_resolve(c)) return rc;
    // the recombination candidates of the tick tracing these rays: every hit
    c->rec_cut = INT64_MAX;
    // the trace indexes the field roots with each ray's field: every index is
    // checked before the pool is touched, and a refused upload leaves no pool
    int32_t maxf = -1;
    for (int64_t i = 0; i < nray; i++) {
        if (rays[i].field < 0 || rays[i].field >= std::max(c->nfield, 1)) {
            c->nray_pool = 0;
            c->ray_field_max = -1;
            PSR_FAIL(c, PSR_ERR_ARG, "ray field index out of range");
        }
        maxf = std::max(maxf, rays[i].field);
    }
    if (nray > c->ray_cap) {
        dfree(c->d_rays);
        PSR_HIP(c, hipMalloc(&c->d_rays, sizeof(psr_ray) * nray));
        c->ray_cap = nray;
    }
    if (nray) PSR_HIP(c, hipMemcpyAsync(c->d_rays, rays, sizeof(psr_ray) * nray, hipMemcpyHostToDevice, c->stream));
    PSR_HIP(c, hipStreamSynchronize(c->stream));
    c->nray_pool = nray;
    c->ray_field_max = maxf;
    return PSR_OK;
}

extern "C" int psr_rays_download(psr_ctx *c, int64_t first, int64_t nray, psr_ray *rays) {
    if (!c || first < 0 || nray < 0 || (nray > 0 && !rays)) return PSR_ERR_ARG;
    if (first + nray > c->nray_pool) PSR_FAIL(c, PSR_ERR_ARG, "psr_rays_download: range outside the pool");
    PSR_HIP(c, hipSetDevice(c->device));
    if (int rc_j = psr_join(c)) return rc_j;
    PSR_HIP(c, hipStreamSynchronize(c->stream));
    if (nray) PSR_HIP(c, hipMemcpy(rays, c->d_rays + first, sizeof(psr_ray) * nray, hipMemcpyDeviceToHost));
    return PSR_OK;
}

// ------------------------------------------------------------ stats
extern "C" int psr_get_stats(psr_ctx *c, psr_stats *out) {
    if (!c || !out) return PSR_ERR_ARG;
    std::vector<DevStats> hs(PSR_NSTAT);
    PSR_HIP(c, hipSetDevice(c->device));
    if (int rc_j = psr_join(c)) return rc_j;
    if (int rc = psr_resolve(c)) return rc;
    PSR_HIP(c, hipStreamSynchronize(c->stream));
    PSR_HIP(c, hipMemcpy(hs.data(), c->d_stats, sizeof(DevStats) * PSR_NSTAT, hipMemcpyDeviceToHost));
    DevStats h = hs[0];
    for (int k = 1; k < PSR_NSTAT; k++) {
        const DevStats &o = hs[k];
        h.total_deposit_count += o.total_deposit_count;
        h.saturated_deposit_count += o.saturated_deposit_count;
        h.active_count += o.active_count;
        h.solver_steps += o.solver_steps;
        h.solver_failures += o.solver_failures;
        h.surrogate_count += o.surrogate_count;
        h.surrogate_fallbacks += o.surrogate_fallbacks;
        h.deposit_processed += o.deposit_processed;
        h.lost_photon_count_sum += o.lost_photon_count_sum;
        h.first_ionization_HI += o.first_ionization_HI;
        h.first_ionization_HeI += o.first_ionization_HeI;
        h.first_ionization_HeII += o.first_ionization_HeII;
        h.secondary_ionization_HI += o.secondary_ionization_HI;
        h.secondary_ionization_HeI += o.secondary_ionization_HeI;
    }
    memset(out, 0, sizeof(*out));
    out->ticks = c->ticks;
    out->total_deposit_count = (int64_t)h.total_deposit_count;
    out->saturated_deposit_count = (int64_t)h.saturated_deposit_count;
    out->active_count = (int64_t)h.active_count;
    out->solver_steps = (int64_t)h.solver_steps;
    out->solver_failures = (int64_t)h.solver_failures;
    out->surrogate_count = (int64_t)h.surrogate_count;
    out->surrogate_fallbacks = (int64_t)h.surrogate_fallbacks;
    out->deposit_processed = (int64_t)h.deposit_processed;
    out->lost_photon_count_sum = h.lost_photon_count_sum;
    out->first_ionization_HI = h.first_ionization_HI;
    out->first_ionization_HeI = h.first_ionization_HeI;
    out->first_ionization_HeII = h.first_ionization_HeII;
    out->secondary_ionization_HI = h.secondary_ionization_HI;
    out->secondary_ionization_HeI = h.secondary_ionization_HeI;
    return PSR_OK;
}

extern "C" int psr_reset_stats(psr_ctx *c) {
    if (!c) return PSR_ERR_ARG;
    PSR_HIP(c, hipSetDevice(c->device));
    if (int rc_j = psr_join(c)) return rc_j;
    if (int rc = psr_resolve(c)) return rc;
    PSR_HIP(c, hipMemsetAsync(c->d_stats, 0, sizeof(DevStats) * PSR_NSTAT, c->stream));
    PSR_HIP(c, hipStreamSynchronize(c->stream));
    c->ticks = 0;
    return PSR_OK;
}
