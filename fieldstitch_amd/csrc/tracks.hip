// Feature tracks from verified match lists (replaces recon.py:14-56, build_tracks).
//
// Nodes are the features of all photos, numbered image by image (node = offsets[img] + feature,
// offsets = exclusive cumulative sum of the per-photo feature counts).  Each verified match
// (i, j, q, t) is an undirected edge offsets[i] + q -- offsets[j] + t.  A track is a connected
// component of the used nodes with at least 2 members and no two members in the same photo
// (recon.py:37-48); tracks are numbered in increasing order of their smallest node, and the
// observations are listed track by track, nodes in increasing order (recon.py:49-56; scipy's
// connected_components labels components in order of their smallest node, and build_tracks
// keeps the node order inside a component).
//
// Kernels, one pass each over the nodes or the edges:
//   k_tr_init      parent[x] = x, counters and flags cleared, the (root, image) table emptied
//   k_tr_union     one block per pair: lock-free union-find, the larger root is linked under
//                  the smaller by compare-and-swap, so every root is its component's minimum
//   k_tr_compress  parent[x] = root of x
//   k_tr_mark      per used node: component size += 1; (root, image) inserted into an open-
//                  addressing table, an insert that finds its key already present marks the
//                  component bad (two features of one photo)
//   k_tr_scan_*    exclusive scans of (kept root ? 1 : 0) and (kept root ? size : 0): the track
//                  number and the first output slot of each kept component
//   k_tr_scatter   used node -> a slot of its track (atomic cursor per track)
//   k_tr_emit      per kept track: insertion sort of its slots by node, then tid / img / feat
#include "fsx_common.h"

namespace {

constexpr int TR_THREADS = 256;
constexpr int TR_SCAN_PER_THREAD = 8;
constexpr int TR_SCAN_CHUNK = TR_THREADS * TR_SCAN_PER_THREAD;
constexpr int TR_TOP_THREADS = 1024;
constexpr unsigned long long TR_EMPTY = ~0ull;

// result slots
constexpr int TR_N_OBS = 0, TR_N_TRACKS = 1, TR_ERR_INDEX = 2, TR_ERR_CAP = 3;

struct TrWs {
    int* parent;        // [n]
    int* size;          // [n] used members of the component (at its root)
    int* cursor;        // [n] scatter cursor per track (at its root)
    int* track;         // [n] track number (at kept roots)
    int* first;         // [n] first output slot (at kept roots)
    uint8_t* used;      // [n]
    uint8_t* bad;       // [n] component holds two features of one photo (at its root)
    int* slots;         // [obs_cap] node per output slot
    unsigned long long* table;   // [table_n] (root << 32 | image), TR_EMPTY when free
    int2* block_sums;   // [n_blocks] per scan chunk
    long long table_n, obs_cap, n_blocks;
};

inline size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

long long pow2_at_least(long long x) {
    long long p = 64;
    while (p < x) p <<= 1;
    return p;
}

// Byte offsets of the arrays inside the workspace for n nodes and `obs_cap` possible
// observations (min(n, 2 * edges)).
size_t tr_layout(long long n, long long obs_cap, TrWs* w, char* base) {
    const long long table_n = pow2_at_least(2 * (obs_cap > 0 ? obs_cap : 1));
    const long long n_blocks = (n + TR_SCAN_CHUNK - 1) / TR_SCAN_CHUNK;
    size_t off = 0;
    auto take = [&](size_t bytes) { size_t o = off; off += align256(bytes); return base ? base + o : nullptr; };
    char* p_parent = take(sizeof(int) * n);
    char* p_size = take(sizeof(int) * n);
    char* p_cursor = take(sizeof(int) * n);
    char* p_track = take(sizeof(int) * n);
    char* p_first = take(sizeof(int) * n);
    char* p_used = take(n);
    char* p_bad = take(n);
    char* p_slots = take(sizeof(int) * (obs_cap > 0 ? obs_cap : 1));
    char* p_table = take(sizeof(unsigned long long) * table_n);
    char* p_bsum = take(sizeof(int2) * (n_blocks > 0 ? n_blocks : 1));
    if (w) {
        w->parent = (int*)p_parent; w->size = (int*)p_size; w->cursor = (int*)p_cursor;
        w->track = (int*)p_track; w->first = (int*)p_first; w->used = (uint8_t*)p_used;
        w->bad = (uint8_t*)p_bad; w->slots = (int*)p_slots; w->table = (unsigned long long*)p_table;
        w->block_sums = (int2*)p_bsum; w->table_n = table_n; w->obs_cap = obs_cap; w->n_blocks = n_blocks;
    }
    return off;
}

__device__ __forceinline__ int ld_parent(const int* p, int x) {
    return __hip_atomic_load(p + x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Root of x with path halving.  parent[x] <= x always holds and parents only move to
// ancestors, so every value read is an ancestor of x in the current forest.
__device__ int tr_find(int* parent, int x) {
    int p = ld_parent(parent, x);
    while (p != x) {
        const int gp = ld_parent(parent, p);
        if (gp != p) __hip_atomic_store(parent + x, gp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        x = gp;
        p = ld_parent(parent, x);
    }
    return x;
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
        w.size[x] = 0;
        w.cursor[x] = 0;
        w.used[x] = 0;
        w.bad[x] = 0;
    }
    for (long long h = (long long)blockIdx.x * blockDim.x + threadIdx.x; h < w.table_n; h += stride)
        w.table[h] = TR_EMPTY;
    if (blockIdx.x == 0 && threadIdx.x < 4) result[threadIdx.x] = 0;
}

__global__ void __launch_bounds__(TR_THREADS) k_tr_union(TrWs w, const long long* offsets, int n_img,
                                                         const int* pair_ij, const long long* pair_start,
                                                         const int* q, const int* t, long long n_edges,
                                                         long long n_nodes, long long* result) {
    const int p = blockIdx.x;
    const int i = pair_ij[2 * p], j = pair_ij[2 * p + 1];
    if (i < 0 || i >= n_img || j < 0 || j >= n_img) {
        if (threadIdx.x == 0) result[TR_ERR_INDEX] = 1;
        return;
    }
    const long long oi = offsets[i], ni = offsets[i + 1] - oi;
    const long long oj = offsets[j], nj = offsets[j + 1] - oj;
    const long long s0 = pair_start[p], s1 = pair_start[p + 1];
    if (s0 < 0 || s1 < s0 || s1 > n_edges) {
        if (threadIdx.x == 0) result[TR_ERR_INDEX] = 1;
        return;
    }
    for (long long k = s0 + threadIdx.x; k < s1; k += TR_THREADS) {
        const int a = q[k], b = t[k];
        if (a < 0 || a >= ni || b < 0 || b >= nj || oi + a < 0 || oi + a >= n_nodes || oj + b < 0 ||
            oj + b >= n_nodes) {
            result[TR_ERR_INDEX] = 1;
            continue;
        }
        int u = (int)(oi + a), v = (int)(oj + b);
        w.used[u] = 1;
        w.used[v] = 1;
        while (true) {
            u = tr_find(w.parent, u);
            v = tr_find(w.parent, v);
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
// node's own thread stored its root).
__global__ void k_tr_compress(TrWs w, long long n) {
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long x = (long long)blockIdx.x * blockDim.x + threadIdx.x; x < n; x += stride) {
        int r = (int)x, p = ld_parent(w.parent, r);
        while (p != r) { r = p; p = ld_parent(w.parent, r); }
        __hip_atomic_store(w.parent + x, r, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

__device__ __forceinline__ unsigned long long tr_hash(unsigned long long key) {
    return (key * 0x9E3779B97F4A7C15ull) ^ (key >> 29);
}

__global__ void k_tr_mark(TrWs w, long long n, const long long* offsets, int n_img) {
    const long long stride = (long long)gridDim.x * blockDim.x;
    const unsigned long long mask = (unsigned long long)w.table_n - 1;
    for (long long x = (long long)blockIdx.x * blockDim.x + threadIdx.x; x < n; x += stride) {
        if (!w.used[x]) continue;
        const int r = w.parent[x];
        atomicAdd(w.size + r, 1);
        const unsigned long long key = ((unsigned long long)r << 32) | (unsigned)tr_image(offsets, n_img, x);
        unsigned long long h = tr_hash(key) & mask;
        // load factor <= 1/2 and each node inserts once: the probe ends at a free slot or its key
        while (true) {
            const unsigned long long old = atomicCAS(w.table + h, TR_EMPTY, key);
            if (old == TR_EMPTY) break;
            if (old == key) { w.bad[r] = 1; break; }
            h = (h + 1) & mask;
        }
    }
}

__device__ __forceinline__ int2 tr_flags(const TrWs& w, long long x) {
    const bool kept = w.parent[x] == (int)x && w.size[x] >= 2 && !w.bad[x];
    return kept ? make_int2(1, w.size[x]) : make_int2(0, 0);
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

__global__ void __launch_bounds__(TR_THREADS) k_tr_scan_blocks(TrWs w, long long n) {
    __shared__ int2 s_wave[TR_THREADS / 64];
    const long long b0 = (long long)blockIdx.x * TR_SCAN_CHUNK + (long long)threadIdx.x * TR_SCAN_PER_THREAD;
    int2 sum = make_int2(0, 0);
    for (int k = 0; k < TR_SCAN_PER_THREAD; ++k) {
        if (b0 + k < n) { const int2 f = tr_flags(w, b0 + k); sum.x += f.x; sum.y += f.y; }
    }
    int2 total;
    tr_block_scan(sum, s_wave, total);
    if (threadIdx.x == 0) w.block_sums[blockIdx.x] = total;
}

__global__ void __launch_bounds__(TR_TOP_THREADS) k_tr_scan_top(TrWs w, long long* result, long long obs_cap) {
    __shared__ int2 s_wave[TR_TOP_THREADS / 64];
    int2 carry = make_int2(0, 0);
    for (long long c = 0; c < w.n_blocks; c += TR_TOP_THREADS) {
        const long long b = c + threadIdx.x;
        const int2 v = b < w.n_blocks ? w.block_sums[b] : make_int2(0, 0);
        int2 total;
        const int2 ex = tr_block_scan(v, s_wave, total);
        if (b < w.n_blocks) w.block_sums[b] = make_int2(carry.x + ex.x, carry.y + ex.y);
        carry.x += total.x;
        carry.y += total.y;
    }
    if (threadIdx.x == 0) {
        result[TR_N_OBS] = carry.y;
        result[TR_N_TRACKS] = carry.x;
        if ((long long)carry.y > obs_cap) result[TR_ERR_CAP] = 1;
    }
}

__global__ void __launch_bounds__(TR_THREADS) k_tr_scan_apply(TrWs w, long long n) {
    __shared__ int2 s_wave[TR_THREADS / 64];
    const long long b0 = (long long)blockIdx.x * TR_SCAN_CHUNK + (long long)threadIdx.x * TR_SCAN_PER_THREAD;
    int2 f[TR_SCAN_PER_THREAD];
    int2 sum = make_int2(0, 0);
    for (int k = 0; k < TR_SCAN_PER_THREAD; ++k) {
        f[k] = b0 + k < n ? tr_flags(w, b0 + k) : make_int2(0, 0);
        sum.x += f[k].x;
        sum.y += f[k].y;
    }
    int2 total;
    int2 ex = tr_block_scan(sum, s_wave, total);
    const int2 base = w.block_sums[blockIdx.x];
    ex.x += base.x;
    ex.y += base.y;
    for (int k = 0; k < TR_SCAN_PER_THREAD; ++k) {
        if (f[k].x) { w.track[b0 + k] = ex.x; w.first[b0 + k] = ex.y; }
        ex.x += f[k].x;
        ex.y += f[k].y;
    }
}

__global__ void k_tr_scatter(TrWs w, long long n) {
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long x = (long long)blockIdx.x * blockDim.x + threadIdx.x; x < n; x += stride) {
        if (!w.used[x]) continue;
        const int r = w.parent[x];
        if (!tr_flags(w, r).x) continue;
        const long long slot = (long long)w.first[r] + atomicAdd(w.cursor + r, 1);
        if (slot < w.obs_cap) w.slots[slot] = (int)x;
    }
}

__global__ void k_tr_emit(TrWs w, long long n, const long long* offsets, int n_img, long long* tid,
                          long long* img, long long* feat) {
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long r = (long long)blockIdx.x * blockDim.x + threadIdx.x; r < n; r += stride) {
        const int2 f = tr_flags(w, r);
        if (!f.x) continue;
        const long long s0 = w.first[r], len = f.y;
        if (s0 + len > w.obs_cap) continue;
        int* s = w.slots + s0;
        for (long long a = 1; a < len; ++a) {      // tracks hold at most one feature per photo
            const int v = s[a];
            long long b = a - 1;
            while (b >= 0 && s[b] > v) { s[b + 1] = s[b]; --b; }
            s[b + 1] = v;
        }
        const long long tr = w.track[r];
        for (long long a = 0; a < len; ++a) {
            const long long x = s[a];
            const int im = tr_image(offsets, n_img, x);
            tid[s0 + a] = tr;
            img[s0 + a] = im;
            feat[s0 + a] = x - offsets[im];
        }
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
    const long long init_n = n_nodes > w.table_n ? n_nodes : w.table_n;
    hipLaunchKernelGGL(k_tr_init, dim3(grid_for(init_n)), dim3(TR_THREADS), 0, st, w, n_nodes, result_dev);
    FSX_LAUNCH_CHECK();
    if (n_pair > 0) {
        hipLaunchKernelGGL(k_tr_union, dim3(n_pair), dim3(TR_THREADS), 0, st, w, offsets_dev, n_img, pair_ij_dev,
                           pair_start_dev, q_dev, t_dev, n_edges, n_nodes, result_dev);
        FSX_LAUNCH_CHECK();
    }
    if (n_nodes == 0) return FSX_OK;
    hipLaunchKernelGGL(k_tr_compress, dim3(grid_for(n_nodes)), dim3(TR_THREADS), 0, st, w, n_nodes);
    FSX_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_tr_mark, dim3(grid_for(n_nodes)), dim3(TR_THREADS), 0, st, w, n_nodes, offsets_dev, n_img);
    FSX_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_tr_scan_blocks, dim3((unsigned)w.n_blocks), dim3(TR_THREADS), 0, st, w, n_nodes);
    FSX_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_tr_scan_top, dim3(1), dim3(TR_TOP_THREADS), 0, st, w, result_dev, cap);
    FSX_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_tr_scan_apply, dim3((unsigned)w.n_blocks), dim3(TR_THREADS), 0, st, w, n_nodes);
    FSX_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_tr_scatter, dim3(grid_for(n_nodes)), dim3(TR_THREADS), 0, st, w, n_nodes);
    FSX_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_tr_emit, dim3(grid_for(n_nodes)), dim3(TR_THREADS), 0, st, w, n_nodes, offsets_dev, n_img,
                       tid_dev, img_dev, feat_dev);
    FSX_LAUNCH_CHECK();
    return FSX_OK;
}

}  // extern "C"
