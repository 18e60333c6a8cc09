// Feature tracks from verified match lists (replaces recon.py:14-56, build_tracks).
//
// Nodes are the features of all photos, numbered image by image (node = offsets[img] + feature,
// offsets = exclusive cumulative sum of the per-photo feature counts).  Each verified match
// (i, j, q, t) is an undirected edge offsets[i] + q -- offsets[j] + t.  A track is a connected
// component of the matched nodes with at least 2 members and no two members in the same photo
// (recon.py:37-48); tracks are numbered in increasing order of their smallest node, and the
// observations are listed track by track, nodes in increasing order (recon.py:49-56; scipy's
// connected_components labels components in order of their smallest node, and build_tracks
// keeps the node order inside a component).
//
// Kernels:
//   k_tr_init         parent[x] = x, counters cleared
//   k_tr_union        1024 matches per block: lock-free union-find, the larger root is linked under
//                     the smaller by compare-and-swap, so every root is its component's minimum
//   k_tr_compress     parent[x] = root of x; the root counts its members, each member keeps
//                     the count it found as its index
//   k_tr_scan_*<0>    over nodes: components with >= 2 members ("candidates") get a
//                     number and a run of slots
//   k_tr_scatter      member of a candidate -> its slot (index from the compress kernel)
//   k_tr_rank         one thread per slot: the member's rank (members below it) places it in
//                     node order; a second member in its photo's node range, or more members
//                     than photos, clears the candidate
//   k_tr_scan_*<1>    over candidates: tracks (no repeated photo) get a number and an output run
//   k_tr_emit         one thread per slot: tid / img / feat of the kept ones, in slot order
#include "fsx_common.h"

namespace {

constexpr int TR_THREADS = 256;
constexpr int TR_SCAN_PER_THREAD = 8;
constexpr int TR_SCAN_CHUNK = TR_THREADS * TR_SCAN_PER_THREAD;
constexpr int TR_TOP_THREADS = 1024;
constexpr int TR_UCHUNK = 1024;      // matches per union block
constexpr int TR_OFF_LDS = 2048;      // photo offsets staged in LDS up to this many entries

// result slots
constexpr int TR_N_OBS = 0, TR_N_TRACKS = 1, TR_ERR_INDEX = 2, TR_ERR_CAP = 3;
// internal counters
constexpr int TR_N_CAND = 0, TR_N_SLOTS = 1;

struct TrWs {
    int* parent;        // [n]
    int* size;          // [n] members of the component (at its root; 1 for a feature without matches)
    int* local;         // [n] index of the node among its component's members (the root: 0)
    int* comp;          // [n] candidate number (at candidate roots)
    int* first;         // [n] first slot (at candidate roots)
    int* slots;         // [obs_cap] member nodes, candidate by candidate
    int* tmp;           // [obs_cap] members in node order, candidate by candidate
    int* cand_root;     // [cand_cap]
    int* cand_ok;       // [cand_cap] 1 = no photo twice
    int* cand_track;    // [cand_cap] track number (kept candidates)
    int* cand_out;      // [cand_cap] first output position (kept candidates)
    int2* block_sums;   // [max(nb_nodes, nb_cand)]
    unsigned long long* counts;  // [4]
    long long obs_cap, cand_cap, nb_nodes, nb_cand;
};

inline size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

// Layout of the workspace for n nodes and obs_cap = min(n, 2 * edges) possible observations
// (each candidate has >= 2 of them, so at most obs_cap / 2 candidates).
size_t tr_layout(long long n, long long obs_cap, TrWs* w, char* base) {
    const long long oc = obs_cap > 0 ? obs_cap : 1;
    const long long cc = obs_cap / 2 > 0 ? obs_cap / 2 : 1;
    const long long nb_nodes = (n + TR_SCAN_CHUNK - 1) / TR_SCAN_CHUNK;
    const long long nb_cand = (cc + TR_SCAN_CHUNK - 1) / TR_SCAN_CHUNK;
    const long long nb = nb_nodes > nb_cand ? nb_nodes : nb_cand;
    size_t off = 0;
    auto take = [&](size_t bytes) { size_t o = off; off += align256(bytes); return base ? base + o : nullptr; };
    char* p[14];
    for (int k = 0; k < 5; ++k) p[k] = take(sizeof(int) * n);
    p[6] = take(sizeof(int) * oc);
    p[7] = take(sizeof(int) * oc);
    for (int k = 8; k < 12; ++k) p[k] = take(sizeof(int) * cc);
    p[12] = take(sizeof(int2) * (nb > 0 ? nb : 1));
    p[13] = take(sizeof(unsigned long long) * 4);
    if (w) {
        w->parent = (int*)p[0]; w->size = (int*)p[1]; w->local = (int*)p[2]; w->comp = (int*)p[3];
        w->first = (int*)p[4]; w->slots = (int*)p[6]; w->tmp = (int*)p[7];
        w->cand_root = (int*)p[8]; w->cand_ok = (int*)p[9]; w->cand_track = (int*)p[10]; w->cand_out = (int*)p[11];
        w->block_sums = (int2*)p[12]; w->counts = (unsigned long long*)p[13];
        w->obs_cap = oc; w->cand_cap = cc; w->nb_nodes = nb_nodes; w->nb_cand = nb_cand;
    }
    return off;
}

__device__ __forceinline__ int ld_parent(const int* p, int x) {
    return __hip_atomic_load(p + x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Roots of a and b by path halving.  parent[x] <= x always holds and parents only move to
// ancestors, so every value read is an ancestor in the current forest.  The two walks advance
// in lockstep, so their loads overlap instead of running one chain after the other.
__device__ void tr_find2(int* parent, int& a, int& b) {
    int pa = ld_parent(parent, a), pb = ld_parent(parent, b);
    while (pa != a || pb != b) {
        const int ga = pa != a ? ld_parent(parent, pa) : pa;
        const int gb = pb != b ? ld_parent(parent, pb) : pb;
        if (pa != a) {
            if (ga != pa) __hip_atomic_store(parent + a, ga, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            a = ga;
            pa = ld_parent(parent, a);
        }
        if (pb != b) {
            if (gb != pb) __hip_atomic_store(parent + b, gb, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            b = gb;
            pb = ld_parent(parent, b);
        }
    }
}

// Image of node x: the last image whose first node is <= x (offsets has n_img + 1 entries).
__device__ __forceinline__ int tr_image(const long long* offsets, int n_img, long long x) {
    int lo = 0, hi = n_img;            // offsets[lo] <= x < offsets[hi]
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (offsets[mid] <= x) lo = mid; else hi = mid;
    }
    return lo;
}

__global__ void k_tr_init(TrWs w, long long n, long long* result) {
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long x = (long long)blockIdx.x * blockDim.x + threadIdx.x; x < n; x += stride) {
        w.parent[x] = (int)x;
        w.size[x] = 1;
    }
    if (blockIdx.x == 0 && threadIdx.x < 4) {
        result[threadIdx.x] = 0;
        w.counts[threadIdx.x] = 0;
    }
}

// Last pair whose first match is <= e among pairs [lo, hi) (pair_start rises; a pair
// without matches shares its start with the next, so the last such pair is the one holding e).
__device__ __forceinline__ int tr_pair_of(const long long* pair_start, int lo, int hi, long long e) {
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (pair_start[mid] <= e) lo = mid; else hi = mid;
    }
    return lo;
}

// TR_UCHUNK consecutive matches per block, whatever pairs they belong to (one block per pair
// left the largest pairs on the critical path).
__global__ void __launch_bounds__(TR_THREADS) k_tr_union(TrWs w, const long long* offsets, int n_img,
                                                         const int* pair_ij, const long long* pair_start,
                                                         int n_pair, const int* q, const int* t, long long n_edges,
                                                         long long n_nodes, long long* result) {
    __shared__ int s_range[2];
    const long long base = (long long)blockIdx.x * TR_UCHUNK;
    const long long last = (base + TR_UCHUNK < n_edges ? base + TR_UCHUNK : n_edges) - 1;
    if (threadIdx.x < 2) s_range[threadIdx.x] = tr_pair_of(pair_start, 0, n_pair, threadIdx.x ? last : base);
    __syncthreads();
    const int plo = s_range[0], phi = s_range[1];
    for (int m = 0; m < TR_UCHUNK / TR_THREADS; ++m) {
        const long long k = base + m * TR_THREADS + threadIdx.x;
        if (k > last) break;
        const int p = tr_pair_of(pair_start, plo, phi + 1, k);
        const int i = pair_ij[2 * p], j = pair_ij[2 * p + 1];
        if (pair_start[p] > k || pair_start[p + 1] <= k || i < 0 || i >= n_img || j < 0 || j >= n_img) {
            result[TR_ERR_INDEX] = 1;
            continue;
        }
        const long long oi = offsets[i], ni = offsets[i + 1] - oi;
        const long long oj = offsets[j], nj = offsets[j + 1] - oj;
        const int a = q[k], b = t[k];
        if (a < 0 || a >= ni || b < 0 || b >= nj || oi + a < 0 || oi + a >= n_nodes || oj + b < 0 ||
            oj + b >= n_nodes) {
            result[TR_ERR_INDEX] = 1;
            continue;
        }
        int u = (int)(oi + a), v = (int)(oj + b);
        while (true) {
            tr_find2(w.parent, u, v);
            if (u == v) break;
            if (u < v) { const int s = u; u = v; v = s; }
            // link the larger root u under v; a failed swap returns u's new parent
            const int old = atomicCAS(w.parent + u, u, v);
            if (old == u) break;
            u = old;
        }
    }
}

// After the unions: every thread stores only roots, so no parent is left pointing below its
// root when the kernel ends (path halving here could store a non-root ancestor after the
// node's own thread stored its root).  Sizes start at 1 and every non-root adds itself to its
// root; the count it finds there is its index among the members.  A feature without matches is never linked: it stays a root of size 1, so the
// components of >= 2 members are exactly those of matched features.
__global__ void k_tr_compress(TrWs w, long long n) {
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long x = (long long)blockIdx.x * blockDim.x + threadIdx.x; x < n; x += stride) {
        int r = (int)x, p = ld_parent(w.parent, r);
        while (p != r) { r = p; p = ld_parent(w.parent, r); }
        __hip_atomic_store(w.parent + x, r, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        w.local[x] = r != (int)x ? atomicAdd(w.size + r, 1) : 0;
    }
}

// Candidates to scan in the second pass: none when a capacity was exceeded (the sort and
// scatter kernels then skipped their work and the host reports the capacity error).
__device__ __forceinline__ long long tr_live_candidates(const TrWs& w) {
    const unsigned long long nc = w.counts[TR_N_CAND], ns = w.counts[TR_N_SLOTS];
    return (nc <= (unsigned long long)w.cand_cap && ns <= (unsigned long long)w.obs_cap) ? (long long)nc : 0;
}

// Scan inputs: MODE 0 over nodes (candidate roots: 1, members), MODE 1 over candidates
// (kept tracks: 1, members).
template <int MODE>
__device__ __forceinline__ int2 tr_flags(const TrWs& w, long long x, long long n_cand) {
    if (MODE == 0) {
        const int s = w.size[x];
        return (w.parent[x] == (int)x && s >= 2) ? make_int2(1, s) : make_int2(0, 0);
    }
    if (x >= n_cand || !w.cand_ok[x]) return make_int2(0, 0);
    return make_int2(1, w.size[w.cand_root[x]]);
}

// Block-wide exclusive scan of one int2 per thread; returns the prefix, `total` gets the sum.
__device__ int2 tr_block_scan(int2 v, int2* s_wave, int2& total) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
    int2 inc = v;
    for (int d = 1; d < 64; d <<= 1) {
        const int ax = __shfl_up(inc.x, d), ay = __shfl_up(inc.y, d);
        if (lane >= d) { inc.x += ax; inc.y += ay; }
    }
    if (lane == 63) s_wave[wid] = inc;
    __syncthreads();
    int2 base = make_int2(0, 0);
    total = make_int2(0, 0);
    for (int k = 0; k < nw; ++k) {
        const int2 s = s_wave[k];
        if (k < wid) { base.x += s.x; base.y += s.y; }
        total.x += s.x; total.y += s.y;
    }
    __syncthreads();
    return make_int2(base.x + inc.x - v.x, base.y + inc.y - v.y);
}

template <int MODE>
__global__ void __launch_bounds__(TR_THREADS) k_tr_scan_blocks(TrWs w, long long n) {
    __shared__ int2 s_wave[TR_THREADS / 64];
    const long long n_cand = MODE ? tr_live_candidates(w) : 0;
    const long long b0 = (long long)blockIdx.x * TR_SCAN_CHUNK + (long long)threadIdx.x * TR_SCAN_PER_THREAD;
    int2 sum = make_int2(0, 0);
    for (int k = 0; k < TR_SCAN_PER_THREAD; ++k) {
        if (b0 + k < n) { const int2 f = tr_flags<MODE>(w, b0 + k, n_cand); sum.x += f.x; sum.y += f.y; }
    }
    int2 total;
    tr_block_scan(sum, s_wave, total);
    if (threadIdx.x == 0) w.block_sums[blockIdx.x] = total;
}

template <int MODE>
__global__ void __launch_bounds__(TR_TOP_THREADS) k_tr_scan_top(TrWs w, long long nb, long long* result) {
    __shared__ int2 s_wave[TR_TOP_THREADS / 64];
    int2 carry = make_int2(0, 0);
    for (long long c = 0; c < nb; c += TR_TOP_THREADS) {
        const long long b = c + threadIdx.x;
        const int2 v = b < nb ? w.block_sums[b] : make_int2(0, 0);
        int2 total;
        const int2 ex = tr_block_scan(v, s_wave, total);
        if (b < nb) w.block_sums[b] = make_int2(carry.x + ex.x, carry.y + ex.y);
        carry.x += total.x;
        carry.y += total.y;
    }
    if (threadIdx.x == 0) {
        if (MODE == 0) {
            w.counts[TR_N_CAND] = (unsigned long long)carry.x;
            w.counts[TR_N_SLOTS] = (unsigned long long)carry.y;
            if ((long long)carry.y > w.obs_cap || (long long)carry.x > w.cand_cap) result[TR_ERR_CAP] = 1;
        } else {
            result[TR_N_OBS] = carry.y;
            result[TR_N_TRACKS] = carry.x;
        }
    }
}

template <int MODE>
__global__ void __launch_bounds__(TR_THREADS) k_tr_scan_apply(TrWs w, long long n) {
    __shared__ int2 s_wave[TR_THREADS / 64];
    const long long n_cand = MODE ? tr_live_candidates(w) : 0;
    const long long b0 = (long long)blockIdx.x * TR_SCAN_CHUNK + (long long)threadIdx.x * TR_SCAN_PER_THREAD;
    int2 f[TR_SCAN_PER_THREAD];
    int2 sum = make_int2(0, 0);
    for (int k = 0; k < TR_SCAN_PER_THREAD; ++k) {
        f[k] = b0 + k < n ? tr_flags<MODE>(w, b0 + k, n_cand) : make_int2(0, 0);
        sum.x += f[k].x;
        sum.y += f[k].y;
    }
    int2 total;
    int2 ex = tr_block_scan(sum, s_wave, total);
    const int2 base = w.block_sums[blockIdx.x];
    ex.x += base.x;
    ex.y += base.y;
    for (int k = 0; k < TR_SCAN_PER_THREAD; ++k) {
        if (f[k].x) {
            if (MODE == 0) {
                if (ex.x < w.cand_cap) {
                    w.comp[b0 + k] = ex.x;
                    w.first[b0 + k] = ex.y;
                    w.cand_root[ex.x] = (int)(b0 + k);
                    w.cand_ok[ex.x] = 1;
                }
            } else {
                w.cand_track[b0 + k] = ex.x;
                w.cand_out[b0 + k] = ex.y;
            }
        }
        ex.x += f[k].x;
        ex.y += f[k].y;
    }
}

__global__ void k_tr_scatter(TrWs w, long long n) {
    if (w.counts[TR_N_SLOTS] > (unsigned long long)w.obs_cap || w.counts[TR_N_CAND] > (unsigned long long)w.cand_cap)
        return;
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long x = (long long)blockIdx.x * blockDim.x + threadIdx.x; x < n; x += stride) {
        const int r = w.parent[x];
        if (w.size[r] < 2) continue;
        w.slots[w.first[r] + w.local[x]] = (int)x;
    }
}

// Photo offsets in LDS when they fit, else the global table.
__device__ __forceinline__ const long long* tr_stage_offsets(long long* s_off, const long long* offsets, int n_img) {
    if (n_img + 1 > TR_OFF_LDS) return offsets;
    for (int k = threadIdx.x; k <= n_img; k += blockDim.x) s_off[k] = offsets[k];
    __syncthreads();
    return s_off;
}

// One thread per slot.  Its member x gets rank = members below x and counts the members in
// x's photo's node range; the sorted run goes to tmp, and a second member in one photo clears
// the candidate.  A candidate with more members than photos repeats a photo and is skipped.
__global__ void __launch_bounds__(TR_THREADS) k_tr_rank(TrWs w, const long long* offsets, int n_img) {
    __shared__ long long s_off[TR_OFF_LDS];
    const long long* off = tr_stage_offsets(s_off, offsets, n_img);
    const long long n_slots = (long long)w.counts[TR_N_SLOTS];
    if (n_slots > w.obs_cap || w.counts[TR_N_CAND] > (unsigned long long)w.cand_cap) return;
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < n_slots; k += stride) {
        const int x = w.slots[k];
        const int r = w.parent[x];
        const int len = w.size[r];
        const int c = w.comp[r];
        if (len > n_img) {
            w.cand_ok[c] = 0;
            continue;
        }
        const int* seg = w.slots + w.first[r];
        const int im = tr_image(off, n_img, x);
        const long long lo = off[im], hi = off[im + 1];
        int below = 0, same = 0;
        for (int b = 0; b < len; ++b) {
            const int u = seg[b];
            below += u < x;
            same += (u >= lo) & (u < hi);
        }
        w.tmp[w.first[r] + below] = x;      // node ids are distinct: ranks are a permutation
        if (same > 1) w.cand_ok[c] = 0;
    }
}

__global__ void __launch_bounds__(TR_THREADS) k_tr_emit(TrWs w, const long long* offsets, int n_img, long long* tid,
                                                        long long* img, long long* feat) {
    __shared__ long long s_off[TR_OFF_LDS];
    const long long* off = tr_stage_offsets(s_off, offsets, n_img);
    const long long n_slots = (long long)w.counts[TR_N_SLOTS];
    if (n_slots > w.obs_cap || w.counts[TR_N_CAND] > (unsigned long long)w.cand_cap) return;
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < n_slots; k += stride) {
        const int r = w.parent[w.slots[k]];
        const int c = w.comp[r];
        if (!w.cand_ok[c]) continue;
        const int x = w.tmp[k];              // k's place in the sorted run of candidate c
        const long long pos = (long long)w.cand_out[c] + (k - w.first[r]);
        const int im = tr_image(off, n_img, x);
        tid[pos] = w.cand_track[c];
        img[pos] = im;
        feat[pos] = x - off[im];
    }
}

int grid_for(long long n, int cap = 8192) {
    const long long b = (n + TR_THREADS - 1) / TR_THREADS;
    return (int)(b < 1 ? 1 : (b > cap ? cap : b));
}

}  // namespace

extern "C" {

int fsx_tracks_workspace_bytes(long long n_nodes, long long n_edges, size_t* bytes, long long* obs_capacity) {
    FSX_REQUIRE(n_nodes >= 0 && n_nodes < (1ll << 31) && n_edges >= 0 && bytes,
                "fsx_tracks_workspace_bytes: bad n_nodes=%lld n_edges=%lld", n_nodes, n_edges);
    const long long cap = n_nodes < 2 * n_edges ? n_nodes : 2 * n_edges;
    *bytes = tr_layout(n_nodes, cap, nullptr, nullptr);
    if (obs_capacity) *obs_capacity = cap;
    return FSX_OK;
}

int fsx_build_tracks(int n_img, const long long* offsets_dev, long long n_nodes, int n_pair, const int* pair_ij_dev,
                     const long long* pair_start_dev, const int* q_dev, const int* t_dev, long long n_edges,
                     void* ws_dev, size_t ws_bytes, long long out_capacity, long long* tid_dev, long long* img_dev,
                     long long* feat_dev, long long* result_dev, void* stream) {
    FSX_REQUIRE(n_img >= 1 && n_pair >= 0 && n_edges >= 0 && n_nodes >= 0 && n_nodes < (1ll << 31),
                "fsx_build_tracks: bad n_img=%d n_pair=%d n_edges=%lld n_nodes=%lld", n_img, n_pair, n_edges, n_nodes);
    FSX_REQUIRE(offsets_dev && result_dev && ws_dev && (n_pair == 0 || (pair_ij_dev && pair_start_dev)) &&
                    (n_edges == 0 || (q_dev && t_dev)),
                "fsx_build_tracks: null pointer");
    const long long cap = n_nodes < 2 * n_edges ? n_nodes : 2 * n_edges;
    FSX_REQUIRE(out_capacity >= cap && (cap == 0 || (tid_dev && img_dev && feat_dev)),
                "fsx_build_tracks: output capacity %lld < %lld", out_capacity, cap);
    TrWs w;
    const size_t need = tr_layout(n_nodes, cap, &w, (char*)ws_dev);
    FSX_REQUIRE(ws_bytes >= need, "fsx_build_tracks: workspace %zu < %zu bytes", ws_bytes, need);
    hipStream_t st = (hipStream_t)stream;
    hipLaunchKernelGGL(k_tr_init, dim3(grid_for(n_nodes)), dim3(TR_THREADS), 0, st, w, n_nodes, result_dev);
    FSX_LAUNCH_CHECK();
    if (n_pair > 0 && n_edges > 0) {
        hipLaunchKernelGGL(k_tr_union, dim3((unsigned)((n_edges + TR_UCHUNK - 1) / TR_UCHUNK)), dim3(TR_THREADS), 0, st,
                           w, offsets_dev, n_img, pair_ij_dev, pair_start_dev, n_pair, q_dev, t_dev, n_edges, n_nodes,
                           result_dev);
        FSX_LAUNCH_CHECK();
    }
    if (n_nodes == 0) return FSX_OK;
    hipLaunchKernelGGL(k_tr_compress, dim3(grid_for(n_nodes)), dim3(TR_THREADS), 0, st, w, n_nodes);
    FSX_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_tr_scan_blocks<0>, dim3((unsigned)w.nb_nodes), dim3(TR_THREADS), 0, st, w, n_nodes);
    FSX_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_tr_scan_top<0>, dim3(1), dim3(TR_TOP_THREADS), 0, st, w, w.nb_nodes, result_dev);
    FSX_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_tr_scan_apply<0>, dim3((unsigned)w.nb_nodes), dim3(TR_THREADS), 0, st, w, n_nodes);
    FSX_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_tr_scatter, dim3(grid_for(n_nodes)), dim3(TR_THREADS), 0, st, w, n_nodes);
    FSX_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_tr_rank, dim3(grid_for(w.obs_cap)), dim3(TR_THREADS), 0, st, w, offsets_dev, n_img);
    FSX_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_tr_scan_blocks<1>, dim3((unsigned)w.nb_cand), dim3(TR_THREADS), 0, st, w, w.cand_cap);
    FSX_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_tr_scan_top<1>, dim3(1), dim3(TR_TOP_THREADS), 0, st, w, w.nb_cand, result_dev);
    FSX_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_tr_scan_apply<1>, dim3((unsigned)w.nb_cand), dim3(TR_THREADS), 0, st, w, w.cand_cap);
    FSX_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_tr_emit, dim3(grid_for(w.obs_cap)), dim3(TR_THREADS), 0, st, w, offsets_dev, n_img, tid_dev,
                       img_dev, feat_dev);
    FSX_LAUNCH_CHECK();
    return FSX_OK;
}

}  // extern "C"
