// marxs_analysis.hip -- per-order figures of merit of a traced photon list
// on the GPU (MI355X, gfx950).
//
// The reference analyses every tolerancing variant on the host:
// effectivearea_from_photonlist (marxs/analysis/gratings.py:333-363) sums the
// probability of each grating order, resolvingpower_from_photonlist (215-270)
// takes astropy's sigma_clipped_stats of the dispersion coordinate of each
// order (sigma 3, at most 5 rounds, median centre, std with ddof 0: drop
// non-finite values, then repeat {keep values within median -/+ 3 std,
// inclusive} until a round removes nothing).
//
// Here the photon list stays in HBM:
//   1. count   one wave per chunk of photons: per order, the summed
//              probability, the photons counted (flag MX_OS_FILTER: only
//              finite value and probability > 0) and those with a finite
//              value, each lane in its own LDS slot (no atomics: the sums
//              have a fixed order for a given photon count);
//   2. scan    per order (one workgroup each), chunk offsets into one array
//              of segments (one segment per order) and the totals;
//   3. scatter the finite values of each order into its segment;
//   4. sort    all values by value, then (stably) by order: two device-wide
//              hipCUB radix sorts leave every segment sorted (a segmented
//              sort would give each of the few, large segments one
//              workgroup);
//   5. clip    one workgroup per order: on a sorted segment the kept set of
//              every round is a contiguous range, so a round is a median
//              read, two block reductions (mean, then squared deviations)
//              and two binary searches for the new ends.
// The library allocates nothing: all scratch is the caller's workspace.

#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>
#include <mutex>
#include <stdint.h>
#include <stdio.h>
#include <math.h>

#include "marxs_hip.h"

__attribute__((visibility("hidden"))) void mx_set_error(const char *msg);

#define MX_OS_WAVE 64          // work-items per counting / scatter workgroup
#define MX_OS_CLIP 1024        // work-items per clipping workgroup
#define MX_OS_UNROLL 8         // independent loads in flight per work-item (clipping sums)
#define MX_OS_LOADS 8          // photons loaded ahead per lane (counting, scatter)
#define MX_OS_CHUNK 4096       // photons per counting workgroup (at least)
#define MX_OS_MAX_BLOCKS 16384
#define MX_OS_SCAN 256         // work-items per scan workgroup (one per order)

static int64_t os_blocks(int64_t n) {
    int64_t nb = (n + MX_OS_CHUNK - 1) / MX_OS_CHUNK;
    if (nb < 1) nb = 1;
    if (nb > MX_OS_MAX_BLOCKS) nb = MX_OS_MAX_BLOCKS;
    return nb;
}

static size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

// Workspace layout (bytes, each part 256-aligned).
struct OsLayout {
    size_t cnt_b, nval_b, psum_b, off_b, tot, seg, vals, sorted, ord, ord2, clip, part0, part1, temp,
        temp_bytes, total;
};

#define MX_OS_CW 16            // workgroups per order in the clipping sums

// Clipping state of one order between the launches of the rounds.
struct ClipState {
    int64_t lo, hi;            // current range of the sorted segment
    double med, mean, sd;      // statistics of that range when `have`
    int32_t done, have;        // no further round; statistics final
};

// temporary storage of the two radix sorts (by value with the order index
// as payload, then by order index with the value as payload)
static int os_sort_temp_bytes(int64_t n, size_t &bytes) {
    size_t b1 = 0, b2 = 0;
    if (hipcub::DeviceRadixSort::SortPairs(nullptr, b1, (const double *)nullptr, (double *)nullptr,
                                           (const uint8_t *)nullptr, (uint8_t *)nullptr, (int)n, 0, 64,
                                           (hipStream_t)0) != hipSuccess)
        return 1;
    if (hipcub::DeviceRadixSort::SortPairs(nullptr, b2, (const uint8_t *)nullptr, (uint8_t *)nullptr,
                                           (const double *)nullptr, (double *)nullptr, (int)n, 0, 8,
                                           (hipStream_t)0) != hipSuccess)
        return 1;
    bytes = b1 > b2 ? b1 : b2;
    return 0;
}

static int os_layout(int64_t n, int no, bool values, OsLayout &L) {
    const int64_t nb = os_blocks(n);
    size_t o = 0;
    L.cnt_b = o;  o += align256(sizeof(int64_t) * nb * no);
    L.nval_b = o; o += align256(sizeof(int64_t) * nb * no);
    L.psum_b = o; o += align256(sizeof(double) * nb * no);
    L.off_b = o;  o += align256(sizeof(int64_t) * nb * no);
    L.tot = o;    o += align256(sizeof(int64_t) * no);
    L.seg = o;    o += align256(sizeof(int32_t) * 2 * no);
    L.vals = o;   o += values ? align256(sizeof(double) * (n > 0 ? n : 1)) : 0;
    L.sorted = o; o += values ? align256(sizeof(double) * (n > 0 ? n : 1)) : 0;
    L.ord = o;    o += values ? align256(n > 0 ? n : 1) : 0;
    L.ord2 = o;   o += values ? align256(n > 0 ? n : 1) : 0;
    L.clip = o;   o += values ? align256(sizeof(ClipState) * no) : 0;
    L.part0 = o;  o += values ? align256(sizeof(double) * no * MX_OS_CW) : 0;
    L.part1 = o;  o += values ? align256(sizeof(double) * no * MX_OS_CW) : 0;
    L.temp = o;
    L.temp_bytes = 0;
    if (values && os_sort_temp_bytes(n, L.temp_bytes)) return 1;
    o += align256(L.temp_bytes > 0 ? L.temp_bytes : 1);
    L.total = o;
    return 0;
}

// index of order o in the list (linear: at most MX_MAX_ORDERS entries), -1
__device__ __forceinline__ int order_index(const int *ords, int no, int o) {
    int k = -1;
    for (int j = 0; j < no; ++j)
        if (ords[j] == o) { k = j; break; }
    return k;
}

// Order number -> index through a table of MX_OS_LUT entries in LDS when the
// orders span fewer values (the usual case), else the linear search.  Built
// by the one wave of a counting / scatter workgroup.
#define MX_OS_LUT 256
struct OrderLut {
    const int *t;
    const int *ords;
    int lo, no;
    bool ok;
    __device__ __forceinline__ int operator()(int o) const {
        if (!ok) return order_index(ords, no, o);
        const int64_t d = (int64_t)o - lo;
        const int k = (d >= 0 && d < MX_OS_LUT) ? t[d] : MX_MAX_ORDERS;
        return k < MX_MAX_ORDERS ? k : -1;
    }
};

__device__ __forceinline__ OrderLut build_lut(const MxOrderStats &a, int *ords, int *lut) {
    const int no = a.n_orders, lane = threadIdx.x;
    int mn = 0x7fffffff, mx = INT32_MIN;
    for (int k = lane; k < no; k += MX_OS_WAVE) {
        const int o = a.orders[k];
        ords[k] = o;
        mn = o < mn ? o : mn;
        mx = o > mx ? o : mx;
    }
    for (int w = 32; w > 0; w >>= 1) {
        const int omn = __shfl_xor(mn, w, MX_OS_WAVE), omx = __shfl_xor(mx, w, MX_OS_WAVE);
        mn = omn < mn ? omn : mn;
        mx = omx > mx ? omx : mx;
    }
    OrderLut L;
    L.t = lut;
    L.ords = ords;
    L.lo = mn;
    L.no = no;
    L.ok = (int64_t)mx - (int64_t)mn < MX_OS_LUT;
    for (int j = lane; j < MX_OS_LUT; j += MX_OS_WAVE) lut[j] = MX_MAX_ORDERS;
    __syncthreads();
    // the first index of an order number, as the linear search finds it
    if (L.ok)
        for (int k = lane; k < no; k += MX_OS_WAVE) atomicMin(&lut[(int64_t)a.orders[k] - mn], k);
    __syncthreads();
    return L;
}


// Pass 1: per chunk and order, probability sum, counted photons, finite
// values.  One wave per workgroup; lane l accumulates into slot [k][l].
__global__ __launch_bounds__(MX_OS_WAVE) void mx_os_count_kernel(MxOrderStats a, int64_t chunk,
                                                                  int64_t *cnt_b, int64_t *nval_b,
                                                                  double *psum_b) {
    extern __shared__ __align__(16) unsigned char sm[];
    const int no = a.n_orders, lane = threadIdx.x;
    double *sp = (double *)sm;                          // [no][64]
    int *sc = (int *)(sp + no * MX_OS_WAVE);            // [no][64]
    int *sv = sc + no * MX_OS_WAVE;                     // [no][64]
    int *ords = sv + no * MX_OS_WAVE;                   // [no]
    __shared__ int lut[MX_OS_LUT];
    for (int k = 0; k < no; ++k) {
        sp[k * MX_OS_WAVE + lane] = 0.0;
        sc[k * MX_OS_WAVE + lane] = 0;
        sv[k * MX_OS_WAVE + lane] = 0;
    }
    const OrderLut index_of = build_lut(a, ords, lut);
    const int64_t b0 = (int64_t)blockIdx.x * chunk;
    int64_t b1 = b0 + chunk;
    if (b1 > a.n) b1 = a.n;
    const bool filter = (a.flags & MX_OS_FILTER) != 0;
    // MX_OS_LOADS photons of the lane are loaded before any is counted (the
    // loads overlap); they are counted in index order, as one at a time
    for (int64_t i0 = b0 + lane; i0 < b1; i0 += (int64_t)MX_OS_WAVE * MX_OS_LOADS) {
        int ord[MX_OS_LOADS];
        double pr[MX_OS_LOADS], val[MX_OS_LOADS];
#pragma unroll
        for (int u = 0; u < MX_OS_LOADS; ++u) {
            const int64_t i = i0 + (int64_t)u * MX_OS_WAVE;
            const bool in = i < b1;
            ord[u] = in ? a.order[i] : INT32_MIN;
            pr[u] = in ? a.prob[i] : 0.0;
            val[u] = (in && a.value) ? a.value[i] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < MX_OS_LOADS; ++u) {
            if (i0 + (int64_t)u * MX_OS_WAVE >= b1) break;
            const int k = index_of(ord[u]);
            if (k < 0) continue;
            sp[k * MX_OS_WAVE + lane] += pr[u];
            if (!filter || (isfinite(val[u]) && pr[u] > 0.0)) {
                sc[k * MX_OS_WAVE + lane] += 1;
                if (a.value && isfinite(val[u])) sv[k * MX_OS_WAVE + lane] += 1;
            }
        }
    }
    __syncthreads();
    for (int k = lane; k < no; k += MX_OS_WAVE) {
        double s = 0.0;
        int64_t c = 0, v = 0;
        for (int l = 0; l < MX_OS_WAVE; ++l) {
            s += sp[k * MX_OS_WAVE + l];
            c += sc[k * MX_OS_WAVE + l];
            v += sv[k * MX_OS_WAVE + l];
        }
        const size_t o = (size_t)blockIdx.x * no + k;
        psum_b[o] = s;
        cnt_b[o] = c;
        nval_b[o] = v;
    }
}

// Pass 2: one workgroup per order.  Work-item t takes a contiguous run of
// chunks; the run totals are scanned in order by work-item 0, then every
// work-item writes its chunks' offsets.  Sums run in a fixed order.
__global__ __launch_bounds__(MX_OS_SCAN) void mx_os_scan_kernel(MxOrderStats a, int64_t nb,
                                                                 const int64_t *cnt_b, const int64_t *nval_b,
                                                                 const double *psum_b, int64_t *off_b,
                                                                 int64_t *tot) {
    __shared__ int64_t sv[MX_OS_SCAN], sc[MX_OS_SCAN];
    __shared__ double ss[MX_OS_SCAN];
    const int no = a.n_orders, k = blockIdx.x, t = threadIdx.x;
    const int64_t per = (nb + MX_OS_SCAN - 1) / MX_OS_SCAN;
    const int64_t b0 = (int64_t)t * per;
    const int64_t b1 = b0 + per < nb ? b0 + per : nb;
    int64_t v = 0, c = 0;
    double s = 0.0;
    for (int64_t b = b0; b < b1; ++b) {
        const size_t o = (size_t)b * no + k;
        v += nval_b[o];
        c += cnt_b[o];
        s += psum_b[o];
    }
    sv[t] = v;
    sc[t] = c;
    ss[t] = s;
    __syncthreads();
    if (t == 0) {
        int64_t run = 0, cc = 0;
        double ps = 0.0;
        for (int j = 0; j < MX_OS_SCAN; ++j) {
            const int64_t x = sv[j];
            sv[j] = run;
            run += x;
            cc += sc[j];
            ps += ss[j];
        }
        a.prob_sum[k] = ps;
        a.count[k] = cc;
        tot[k] = run;
    }
    __syncthreads();
    int64_t run = sv[t];
    for (int64_t b = b0; b < b1; ++b) {
        const size_t o = (size_t)b * no + k;
        off_b[o] = run;
        run += nval_b[o];
    }
}

// Segment bounds: order k's values start after those of orders < k.
__global__ void mx_os_seg_kernel(int no, const int64_t *tot, int32_t *seg) {
    const int k = threadIdx.x;
    if (k < no) {
        int64_t start = 0;
        for (int j = 0; j < k; ++j) start += tot[j];
        seg[k] = (int32_t)start;
        seg[no + k] = (int32_t)(start + tot[k]);
    }
}

// Pass 3: values of each order into its segment (lane-ordered within a chunk).
__global__ __launch_bounds__(MX_OS_WAVE) void mx_os_scatter_kernel(MxOrderStats a, int64_t chunk,
                                                                    const int64_t *off_b, const int32_t *seg,
                                                                    double *vals, uint8_t *ordk) {
    extern __shared__ __align__(16) unsigned char sm[];
    const int no = a.n_orders, lane = threadIdx.x;
    int64_t *pos = (int64_t *)sm;                       // [no][64]
    int *ords = (int *)(pos + no * MX_OS_WAVE);         // [no]
    __shared__ int lut[MX_OS_LUT];
    for (int k = 0; k < no; ++k) pos[k * MX_OS_WAVE + lane] = 0;
    const OrderLut index_of = build_lut(a, ords, lut);
    const int64_t b0 = (int64_t)blockIdx.x * chunk;
    int64_t b1 = b0 + chunk;
    if (b1 > a.n) b1 = a.n;
    const bool filter = (a.flags & MX_OS_FILTER) != 0;
    for (int64_t i0 = b0 + lane; i0 < b1; i0 += (int64_t)MX_OS_WAVE * MX_OS_LOADS) {
        int ord[MX_OS_LOADS];
        double pr[MX_OS_LOADS], val[MX_OS_LOADS];
#pragma unroll
        for (int u = 0; u < MX_OS_LOADS; ++u) {
            const int64_t i = i0 + (int64_t)u * MX_OS_WAVE;
            const bool in = i < b1;
            ord[u] = in ? a.order[i] : INT32_MIN;
            pr[u] = (in && filter) ? a.prob[i] : 1.0;
            val[u] = in ? a.value[i] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < MX_OS_LOADS; ++u) {
            if (i0 + (int64_t)u * MX_OS_WAVE >= b1) break;
            const int k = index_of(ord[u]);
            if (k >= 0 && isfinite(val[u]) && (!filter || pr[u] > 0.0)) pos[k * MX_OS_WAVE + lane] += 1;
        }
    }
    __syncthreads();
    // exclusive prefix over lanes, plus the chunk and segment offsets
    for (int k = lane; k < no; k += MX_OS_WAVE) {
        int64_t run = (int64_t)seg[k] + off_b[(size_t)blockIdx.x * no + k];
        for (int l = 0; l < MX_OS_WAVE; ++l) {
            const int64_t c = pos[k * MX_OS_WAVE + l];
            pos[k * MX_OS_WAVE + l] = run;
            run += c;
        }
    }
    __syncthreads();
    for (int64_t i0 = b0 + lane; i0 < b1; i0 += (int64_t)MX_OS_WAVE * MX_OS_LOADS) {
        int ord[MX_OS_LOADS];
        double pr[MX_OS_LOADS], val[MX_OS_LOADS];
#pragma unroll
        for (int u = 0; u < MX_OS_LOADS; ++u) {
            const int64_t i = i0 + (int64_t)u * MX_OS_WAVE;
            const bool in = i < b1;
            ord[u] = in ? a.order[i] : INT32_MIN;
            pr[u] = (in && filter) ? a.prob[i] : 1.0;
            val[u] = in ? a.value[i] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < MX_OS_LOADS; ++u) {
            if (i0 + (int64_t)u * MX_OS_WAVE >= b1) break;
            const int k = index_of(ord[u]);
            if (k >= 0 && isfinite(val[u]) && (!filter || pr[u] > 0.0)) {
                const int64_t p = pos[k * MX_OS_WAVE + lane];
                pos[k * MX_OS_WAVE + lane] = p + 1;
                vals[p] = val[u];
                ordk[p] = (uint8_t)k;
            }
        }
    }
}

// Items [end of the last segment, n) of the sort input: value +inf, order
// index 255 (past any order), so that both sorts leave them at the end.
__global__ void mx_os_tail_kernel(const int32_t *last_end, int64_t n, double *vals, uint8_t *ordk) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && i >= (int64_t)last_end[0]) {
        vals[i] = __builtin_inf();
        ordk[i] = 255;
    }
}

// sum of f(x[j]) over [lo, hi) by the whole workgroup (fixed order: lane
// strides, then a tree in LDS); every work-item gets the result.
template <typename F>
__device__ __forceinline__ double block_sum(const double *x, int64_t lo, int64_t hi, double *red, F f) {
    // MX_OS_UNROLL independent accumulators per work-item, each over its own
    // stride of the range, added in a fixed order
    double acc[MX_OS_UNROLL];
#pragma unroll
    for (int u = 0; u < MX_OS_UNROLL; ++u) acc[u] = 0.0;
    const int64_t step = (int64_t)MX_OS_CLIP * MX_OS_UNROLL;
    int64_t j = lo + threadIdx.x;
    for (; j + (MX_OS_UNROLL - 1) * MX_OS_CLIP < hi; j += step) {
        double v[MX_OS_UNROLL];
#pragma unroll
        for (int u = 0; u < MX_OS_UNROLL; ++u) v[u] = x[j + (int64_t)u * MX_OS_CLIP];
#pragma unroll
        for (int u = 0; u < MX_OS_UNROLL; ++u) acc[u] += f(v[u]);
    }
    for (; j < hi; j += MX_OS_CLIP) acc[0] += f(x[j]);
    double s = 0.0;
#pragma unroll
    for (int u = 0; u < MX_OS_UNROLL; ++u) s += acc[u];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = MX_OS_CLIP / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    const double r = red[0];
    __syncthreads();
    return r;
}

// First index j in [lo, hi) with pred(x[j]) (pred false, then true, along
// the sorted range), or hi; by the whole workgroup: each round every
// work-item tests one of MX_OS_CLIP evenly spaced probes and the range
// shrinks to the gap before the first true probe (the same answer as a
// binary search, in log_1024 rounds of parallel loads instead of log_2
// dependent ones).
template <typename F>
__device__ __forceinline__ int64_t block_first(const double *x, int64_t lo, int64_t hi, int *s_t, F pred) {
    for (;;) {
        const int64_t n = hi - lo;
        if (n <= 0) return lo;
        const int64_t step = n <= MX_OS_CLIP ? 1 : (n + MX_OS_CLIP - 1) / MX_OS_CLIP;
        const int64_t j = lo + (int64_t)threadIdx.x * step;
        if (threadIdx.x == 0) *s_t = MX_OS_CLIP;
        __syncthreads();
        if (j < hi && pred(x[j])) atomicMin(s_t, (int)threadIdx.x);
        __syncthreads();
        const int t = *s_t;
        __syncthreads();
        if (step == 1) return t < MX_OS_CLIP ? lo + t : hi;
        if (t == 0) return lo;
        if (t < MX_OS_CLIP) {
            hi = lo + (int64_t)t * step;                 // pred true here
            lo = lo + (int64_t)(t - 1) * step + 1;       // false at the probe before
        } else {
            lo = lo + (n - 1) / step * step + 1;         // after the last probe (false)
        }
    }
}

__device__ __forceinline__ double median_sorted(const double *x, int64_t lo, int64_t hi) {
    const int64_t m = hi - lo;
    if (m & 1) return x[lo + m / 2];
    return (x[lo + m / 2 - 1] + x[lo + m / 2]) / 2;
}

// Pass 5: sigma clipping of each sorted segment.  Every clipping round keeps
// a contiguous range, so a round is: the sum of the values over the range,
// then the sum of squared deviations from their mean (each by MX_OS_CW
// workgroups per order over consecutive parts of the range, partial sums
// added in part order), then, one workgroup per order, the median, the new
// bounds and the state for the next round.  Rounds are launched up to
// maxiters times; orders that are done return at once.
__global__ void mx_os_clip_init_kernel(int no, const int32_t *seg, ClipState *cs) {
    const int k = threadIdx.x;
    if (k >= no) return;
    ClipState c;
    c.lo = seg[k];
    c.hi = seg[no + k];
    c.med = c.mean = c.sd = __builtin_nan("");
    c.done = c.hi == c.lo;
    c.have = 0;
    cs[k] = c;
}

// pass 0: partial sums of x; pass 1: partial sums of (x - mean)^2, mean from
// pass 0's partials.  final: the passes after the last round, for orders
// whose statistics are not yet final.
__global__ __launch_bounds__(MX_OS_CLIP) void mx_os_clip_sum_kernel(const ClipState *cs, const double *x,
                                                                     double *p0, double *p1, int pass,
                                                                     int final) {
    __shared__ double red[MX_OS_CLIP];
    const int k = blockIdx.x / MX_OS_CW, w = blockIdx.x % MX_OS_CW;
    const ClipState c = cs[k];
    if (final ? (c.have || c.hi == c.lo) : c.done) return;
    const int64_t m = c.hi - c.lo;
    const int64_t b = c.lo + m * w / MX_OS_CW, e = c.lo + m * (w + 1) / MX_OS_CW;
    double s;
    if (pass == 0) {
        s = block_sum(x, b, e, red, [](double v) { return v; });
    } else {
        double t = 0.0;
        for (int q = 0; q < MX_OS_CW; ++q) t += p0[k * MX_OS_CW + q];
        const double mean = t / (double)m;
        s = block_sum(x, b, e, red, [mean](double v) { const double d = v - mean; return d * d; });
    }
    if (threadIdx.x == 0) (pass == 0 ? p0 : p1)[k * MX_OS_CW + w] = s;
}

__device__ __forceinline__ void clip_moments(const double *p0, const double *p1, int k, int64_t m,
                                             double &mean, double &sd) {
    double t0 = 0.0, t1 = 0.0;
    for (int q = 0; q < MX_OS_CW; ++q) t0 += p0[k * MX_OS_CW + q];
    for (int q = 0; q < MX_OS_CW; ++q) t1 += p1[k * MX_OS_CW + q];
    mean = t0 / (double)m;
    sd = sqrt(t1 / (double)m);
}

__global__ __launch_bounds__(MX_OS_CLIP) void mx_os_clip_bounds_kernel(MxOrderStats a, ClipState *cs,
                                                                        const double *x, const double *p0,
                                                                        const double *p1) {
    __shared__ int s_t;
    const int k = blockIdx.x;
    const ClipState c = cs[k];
    if (c.done) return;
    const int64_t m = c.hi - c.lo;
    double mean, sd;
    clip_moments(p0, p1, k, m, mean, sd);
    const double cen = median_sorted(x, c.lo, c.hi);
    const double vmin = cen - sd * a.sigma;
    const double vmax = cen + sd * a.sigma;
    // first index with x >= vmin, then first index with x > vmax
    const int64_t nlo = block_first(x, c.lo, c.hi, &s_t, [vmin](double v) { return v >= vmin; });
    const int64_t nhi = block_first(x, nlo, c.hi, &s_t, [vmax](double v) { return v > vmax; });
    if (threadIdx.x == 0) {
        const int64_t nchanged = m - (nhi - nlo);
        ClipState d;
        d.lo = nlo;
        d.hi = nhi;
        d.med = cen;
        d.mean = mean;
        d.sd = sd;
        d.have = nchanged == 0;
        d.done = nchanged == 0 || nhi == nlo;
        cs[k] = d;
    }
}

__global__ void mx_os_clip_final_kernel(MxOrderStats a, const ClipState *cs, const double *x, const double *p0,
                                        const double *p1) {
    const int k = threadIdx.x;
    if (k >= a.n_orders) return;
    const ClipState c = cs[k];
    const int64_t m = c.hi - c.lo;
    const double nan = __builtin_nan("");
    double mean = nan, med = nan, sd = nan;
    if (c.have) {
        med = c.med;
        mean = c.mean;
        sd = c.sd;
    } else if (m > 0) {
        med = median_sorted(x, c.lo, c.hi);
        clip_moments(p0, p1, k, m, mean, sd);
    }
    a.mean[k] = mean;
    a.median[k] = med;
    a.std[k] = sd;
    a.n_kept[k] = m;
}

extern "C" int64_t mx_order_stats_workspace(int64_t n, int32_t n_orders, int32_t with_values) {
    if (n < 0 || n_orders < 1 || n_orders > MX_MAX_ORDERS) return -1;
    OsLayout L;
    if (os_layout(n, n_orders, with_values != 0, L)) return -1;
    return (int64_t)L.total;
}

extern "C" int mx_order_stats(const MxOrderStats *a, void *stream) {
    if (!a) { mx_set_error("order stats: arguments NULL"); return 1; }
    if (a->n < 0 || a->n > 0x7fffffffLL) { mx_set_error("order stats: photon count outside [0, 2^31)"); return 1; }
    if (a->n_orders < 1 || a->n_orders > MX_MAX_ORDERS) { mx_set_error("order stats: n_orders outside [1, 64]"); return 1; }
    if (!a->orders || !a->prob_sum || !a->count || (a->n > 0 && (!a->order || !a->prob))) {
        mx_set_error("order stats: a required pointer is NULL");
        return 1;
    }
    const bool values = a->value != nullptr;
    if (values && (!a->mean || !a->median || !a->std || !a->n_kept || a->maxiters < 0)) {
        mx_set_error("order stats: statistics outputs NULL or maxiters negative");
        return 1;
    }
    if ((a->flags & MX_OS_FILTER) && !values) {
        mx_set_error("order stats: the finite-value filter needs a value column");
        return 1;
    }
    OsLayout L;
    if (os_layout(a->n, a->n_orders, values, L)) { mx_set_error("order stats: sort size query failed"); return 2; }
    if (!a->workspace || a->workspace_bytes < (int64_t)L.total) {
        mx_set_error("order stats: workspace missing or smaller than mx_order_stats_workspace()");
        return 1;
    }
    hipStream_t st = (hipStream_t)stream;
    unsigned char *ws = (unsigned char *)a->workspace;
    int64_t *cnt_b = (int64_t *)(ws + L.cnt_b);
    int64_t *nval_b = (int64_t *)(ws + L.nval_b);
    double *psum_b = (double *)(ws + L.psum_b);
    int64_t *off_b = (int64_t *)(ws + L.off_b);
    int64_t *tot = (int64_t *)(ws + L.tot);
    int32_t *seg = (int32_t *)(ws + L.seg);
    const int no = a->n_orders;
    const int64_t nb = os_blocks(a->n);
    const int64_t chunk = (a->n + nb - 1) / nb > 0 ? (a->n + nb - 1) / nb : 1;
    const size_t lds1 = (size_t)no * MX_OS_WAVE * (8 + 4 + 4) + (size_t)no * 4;
    hipLaunchKernelGGL(mx_os_count_kernel, dim3((unsigned)nb), dim3(MX_OS_WAVE), lds1, st, *a, chunk,
                       cnt_b, nval_b, psum_b);
    hipLaunchKernelGGL(mx_os_scan_kernel, dim3((unsigned)no), dim3(MX_OS_SCAN), 0, st, *a, nb, cnt_b, nval_b,
                       psum_b, off_b, tot);
    hipLaunchKernelGGL(mx_os_seg_kernel, dim3(1), dim3(MX_OS_WAVE), 0, st, no, (const int64_t *)tot, seg);
    if (values) {
        double *vals = (double *)(ws + L.vals), *sorted = (double *)(ws + L.sorted);
        const size_t lds3 = (size_t)no * MX_OS_WAVE * 8 + (size_t)no * 4;
        uint8_t *ordk = (uint8_t *)(ws + L.ord), *ordk2 = (uint8_t *)(ws + L.ord2);
        hipLaunchKernelGGL(mx_os_scatter_kernel, dim3((unsigned)nb), dim3(MX_OS_WAVE), lds3, st, *a, chunk,
                           off_b, seg, vals, ordk);
        // the segments fill [0, total); total <= n is known on the device
        // only, so the sorts run over n items with the tail set to +inf /
        // an order index past the last, which sorts behind every segment
        hipLaunchKernelGGL(mx_os_tail_kernel, dim3((unsigned)((a->n + 255) / 256 > 0 ? (a->n + 255) / 256 : 1)),
                           dim3(256), 0, st, seg + no + no - 1, a->n, vals, ordk);
        size_t tb = L.temp_bytes;
        if (hipcub::DeviceRadixSort::SortPairs(ws + L.temp, tb, (const double *)vals, sorted,
                                               (const uint8_t *)ordk, ordk2, (int)a->n, 0, 64, st) != hipSuccess) {
            mx_set_error("order stats: sort by value failed");
            return 2;
        }
        tb = L.temp_bytes;
        if (hipcub::DeviceRadixSort::SortPairs(ws + L.temp, tb, (const uint8_t *)ordk2, ordk,
                                               (const double *)sorted, vals, (int)a->n, 0, 8, st) != hipSuccess) {
            mx_set_error("order stats: sort by order failed");
            return 2;
        }
        const double *xs = vals;
        ClipState *cs = (ClipState *)(ws + L.clip);
        double *p0 = (double *)(ws + L.part0), *p1 = (double *)(ws + L.part1);
        hipLaunchKernelGGL(mx_os_clip_init_kernel, dim3(1), dim3(MX_OS_WAVE), 0, st, no, (const int32_t *)seg, cs);
        const dim3 gsum((unsigned)(no * MX_OS_CW));
        for (int it = 0; it < a->maxiters; ++it) {
            hipLaunchKernelGGL(mx_os_clip_sum_kernel, gsum, dim3(MX_OS_CLIP), 0, st, (const ClipState *)cs, xs, p0,
                               p1, 0, 0);
            hipLaunchKernelGGL(mx_os_clip_sum_kernel, gsum, dim3(MX_OS_CLIP), 0, st, (const ClipState *)cs, xs, p0,
                               p1, 1, 0);
            hipLaunchKernelGGL(mx_os_clip_bounds_kernel, dim3((unsigned)no), dim3(MX_OS_CLIP), 0, st, *a, cs, xs,
                               (const double *)p0, (const double *)p1);
        }
        hipLaunchKernelGGL(mx_os_clip_sum_kernel, gsum, dim3(MX_OS_CLIP), 0, st, (const ClipState *)cs, xs, p0, p1,
                           0, 1);
        hipLaunchKernelGGL(mx_os_clip_sum_kernel, gsum, dim3(MX_OS_CLIP), 0, st, (const ClipState *)cs, xs, p0, p1,
                           1, 1);
        hipLaunchKernelGGL(mx_os_clip_final_kernel, dim3(1), dim3(MX_OS_WAVE), 0, st, *a, (const ClipState *)cs,
                           xs, (const double *)p0, (const double *)p1);
    }
    hipError_t err = hipGetLastError();
    if (err != hipSuccess) {
        char buf[256];
        snprintf(buf, sizeof(buf), "order stats: launch failed: %s", hipGetErrorString(err));
        mx_set_error(buf);
        return 2;
    }
    return 0;
}

// ------------------------------------------------ allocation-free variant
// mx_order_stats_static: mx_order_stats with its workspace in `scratch`, a
// file-scope device array of MX_OS_STATIC_BYTES, for callers that cannot
// allocate (tight serving loops, harnesses that forbid hipMalloc; not graph
// capture: the radix sorts clear their pass state with memsets).  Every GPU the
// library runs on holds its own copy until the process ends.  Host code
// never passes the array's name to a kernel or a copy (in host code it is a
// host-side stand-in): the device address of this GPU's copy comes from
// hipGetSymbolAddress on the entry point's first call on that GPU and is
// kept per device.  Calls on one GPU share the array, so calls that overlap
// on one GPU race; it keeps the last call's contents, and every word a call
// reads it has written first (counts, offsets, segments, values, clipping
// state), so nothing needs clearing.
static __device__ __align__(256) unsigned char scratch[MX_OS_STATIC_BYTES];

#define MX_OS_MAX_DEVICES 64
static std::mutex g_scratch_mutex;
static void *g_scratch_dev[MX_OS_MAX_DEVICES];

static int scratch_address(void **out) {
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= MX_OS_MAX_DEVICES) {
        mx_set_error("order stats (static workspace): no current device or device index too large");
        return 2;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_dev[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            char buf[256];
            snprintf(buf, sizeof(buf), "order stats (static workspace): hipGetSymbolAddress failed: %s",
                     hipGetErrorString(e));
            mx_set_error(buf);
            return 2;
        }
        g_scratch_dev[dev] = p;
    }
    *out = g_scratch_dev[dev];
    return 0;
}

extern "C" int mx_order_stats_static(const MxOrderStats *a, void *stream) {
    if (!a) { mx_set_error("order stats: arguments NULL"); return 1; }
    if (a->n < 0 || a->n > 0x7fffffffLL) { mx_set_error("order stats: photon count outside [0, 2^31)"); return 1; }
    if (a->n_orders < 1 || a->n_orders > MX_MAX_ORDERS) { mx_set_error("order stats: n_orders outside [1, 64]"); return 1; }
    OsLayout L;
    if (os_layout(a->n, a->n_orders, a->value != nullptr, L)) {
        mx_set_error("order stats: sort size query failed");
        return 2;
    }
    if ((int64_t)L.total > (int64_t)MX_OS_STATIC_BYTES) {
        char buf[256];
        snprintf(buf, sizeof(buf),
                 "order stats (static workspace): %lld photons need %lld bytes of workspace, the static "
                 "workspace holds %lld (use mx_order_stats with a workspace of your own)",
                 (long long)a->n, (long long)L.total, (long long)MX_OS_STATIC_BYTES);
        mx_set_error(buf);
        return 1;
    }
    void *ws = nullptr;
    if (scratch_address(&ws)) return 2;
    MxOrderStats b = *a;
    b.workspace = ws;
    b.workspace_bytes = (int64_t)MX_OS_STATIC_BYTES;
    return mx_order_stats(&b, stream);
}
