// halo.hip -- device build of the per-tile halo lists (see halo.hpp).
//
// Both lists come from one pattern: emit a 64-bit key (tile << 32 | cell) per
// (cell, slot) pair that qualifies, radix-sort, drop duplicates, and find each
// tile's range by binary search.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <string>
#include <vector>

#include "common.hpp"
#include "device_math.hpp"
#include "halo.hpp"

namespace hcfd {
namespace {

#define HIPCHK(x)                                                                                   \
    do {                                                                                            \
        hipError_t e_ = (x);                                                                        \
        if (e_ != hipSuccess) fail(HCFD_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

using u64 = unsigned long long;
constexpr u64 NO_KEY = ~0ull;
constexpr u64 CLS_REST = 1ull << 30;  // S2 entries that are not ghosts of the tile's own cells
constexpr u64 ID_MASK = (1ull << 30) - 1;

template <typename T>
T* dalloc(size_t n) {
    T* p = nullptr;
    hipError_t e = hipMalloc(&p, sizeof(T) * (n ? n : 1));
    if (e != hipSuccess) fail(HCFD_ERR_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
    return p;
}

inline unsigned nblk(long long n) { return (unsigned)((n + 255) / 256); }

__device__ __forceinline__ long long lower_bound_u64(const u64* a, long long n, u64 x) {
    long long lo = 0, hi = n;
    while (lo < hi) {
        const long long mid = (lo + hi) >> 1;
        if (a[mid] < x) lo = mid + 1; else hi = mid;
    }
    return lo;
}

// Tile of each interior cell (tiles are consecutive ranges tile_start[t]..).
__global__ void k_cell_tiles(int Nint, int ntiles, const int* __restrict__ tile_start, int* __restrict__ ctile) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= Nint) return;
    int lo = 0, hi = ntiles - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (tile_start[mid] <= i) lo = mid; else hi = mid - 1;
    }
    ctile[i] = lo;
}

// H1 candidates: interior neighbours in another tile.
__global__ void k_h1_keys(int Nint, long long P, int nslot, const int* __restrict__ ctile, const int* __restrict__ cnb,
                          u64* __restrict__ keys) {
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (long long)nslot * Nint) return;
    const int k = (int)(q / Nint), i = (int)(q % Nint);
    const int x = cnb[k * P + i];
    u64 key = NO_KEY;
    if (x != SLOT_EMPTY) {
        const int j = x & SLOT_NB_MASK;
        if (j < Nint && ctile[j] != ctile[i]) key = ((u64)ctile[i] << 32) | (u64)j;
    }
    keys[q] = key;
}

// S2 candidates: ghosts of tile cells (class 0), and neighbours of H1 cells
// that are neither in the tile nor in its H1 list (class 1).
__global__ void k_s2_keys(int Nint, long long P, int nslot, const int* __restrict__ ctile, const int* __restrict__ cnb,
                          const u64* __restrict__ h1keys, long long nh1, u64* __restrict__ keys) {
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long n_own = (long long)nslot * Nint;
    if (q >= n_own + (long long)nslot * nh1) return;
    u64 key = NO_KEY;
    if (q < n_own) {
        const int k = (int)(q / Nint), i = (int)(q % Nint);
        const int x = cnb[k * P + i];
        if (x != SLOT_EMPTY) {
            const int j = x & SLOT_NB_MASK;
            if (j >= Nint) key = ((u64)ctile[i] << 32) | (u64)j;
        }
    } else {
        const long long r = q - n_own;
        const int k = (int)(r / nh1);
        const long long e = r % nh1;
        const u64 hk = h1keys[e];
        const u64 t = hk >> 32;
        const int c = (int)(hk & 0xffffffffull);
        const int x = cnb[k * P + c];
        if (x != SLOT_EMPTY) {
            const int j = x & SLOT_NB_MASK;
            bool skip = false;
            if (j < Nint && (u64)ctile[j] == t) skip = true;
            if (!skip && j < Nint) {
                const long long p = lower_bound_u64(h1keys, nh1, (t << 32) | (u64)j);
                skip = p < nh1 && h1keys[p] == ((t << 32) | (u64)j);
            }
            if (!skip) key = (t << 32) | CLS_REST | (u64)j;
        }
    }
    keys[q] = key;
}

__global__ void k_tile_bounds(int ntiles, const u64* __restrict__ keys, long long n, u64 cls, int* __restrict__ off) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t > ntiles) return;
    off[t] = (int)lower_bound_u64(keys, n, ((u64)t << 32) | cls);
}

__global__ void k_key_ids(long long n, const u64* __restrict__ keys, int* __restrict__ ids) {
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) ids[q] = (int)(keys[q] & ID_MASK);
}

// Sorts keys[0, n) and removes duplicates and NO_KEY entries; returns the count.
long long sort_unique(u64*& keys, long long n, hipStream_t st) {
    if (n > 0x7fffffffLL) fail(HCFD_ERR_ARG, "halo key list too long");
    u64* alt = dalloc<u64>(n);
    hipcub::DoubleBuffer<u64> kb(keys, alt);
    size_t bytes = 0;
    HIPCHK(hipcub::DeviceRadixSort::SortKeys(nullptr, bytes, kb, (int)n, 0, 64, st));
    void* tmp = dalloc<char>(bytes);
    HIPCHK(hipcub::DeviceRadixSort::SortKeys(tmp, bytes, kb, (int)n, 0, 64, st));
    HIPCHK(hipStreamSynchronize(st));
    (void)hipFree(tmp);
    u64* sorted = kb.Current();
    u64* other = sorted == keys ? alt : keys;
    int* d_cnt = dalloc<int>(1);
    bytes = 0;
    HIPCHK(hipcub::DeviceSelect::Unique(nullptr, bytes, sorted, other, d_cnt, (int)n, st));
    tmp = dalloc<char>(bytes);
    HIPCHK(hipcub::DeviceSelect::Unique(tmp, bytes, sorted, other, d_cnt, (int)n, st));
    int cnt = 0;
    HIPCHK(hipMemcpyAsync(&cnt, d_cnt, sizeof(int), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    (void)hipFree(tmp);
    (void)hipFree(d_cnt);
    (void)hipFree(sorted);
    keys = other;
    if (cnt > 0) {
        u64 last = 0;
        HIPCHK(hipMemcpy(&last, keys + cnt - 1, sizeof(u64), hipMemcpyDeviceToHost));
        if (last == NO_KEY) cnt--;
    }
    return cnt;
}

}  // namespace

namespace {

// Lists of one tiling (tile_start on the host); per-tile sizes returned.
void build_lists(int Nint, long long P, int nslot, const int* cnb, const std::vector<int>& ts, HaloTables& out,
                 std::vector<int>& nh, std::vector<int>& ns, std::vector<int>& ng, hipStream_t st) {
    const int ntiles = (int)ts.size() - 1;
    out.ntiles = ntiles;
    out.tile_start = dalloc<int>(ntiles + 1);
    HIPCHK(hipMemcpyAsync(out.tile_start, ts.data(), sizeof(int) * (ntiles + 1), hipMemcpyHostToDevice, st));
    int* ctile = dalloc<int>(Nint);
    k_cell_tiles<<<nblk(Nint), 256, 0, st>>>(Nint, ntiles, out.tile_start, ctile);
    const long long n1 = (long long)nslot * Nint;
    // ---- H1 ---------------------------------------------------------------------------
    u64* k1 = dalloc<u64>(n1);
    k_h1_keys<<<nblk(n1), 256, 0, st>>>(Nint, P, nslot, ctile, cnb, k1);
    HIPCHK(hipGetLastError());
    const long long nh1 = sort_unique(k1, n1, st);
    out.nh1 = nh1;
    out.h1_off = dalloc<int>(ntiles + 1);
    out.h1 = dalloc<int>(nh1);
    k_tile_bounds<<<nblk(ntiles + 1), 256, 0, st>>>(ntiles, k1, nh1, 0, out.h1_off);
    if (nh1) k_key_ids<<<nblk(nh1), 256, 0, st>>>(nh1, k1, out.h1);
    // ---- S2 ---------------------------------------------------------------------------
    const long long n2 = n1 + (long long)nslot * nh1;
    u64* k2 = dalloc<u64>(n2);
    k_s2_keys<<<nblk(n2), 256, 0, st>>>(Nint, P, nslot, ctile, cnb, k1, nh1, k2);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(st));
    (void)hipFree(k1);
    (void)hipFree(ctile);
    const long long ns2 = sort_unique(k2, n2, st);
    out.ns2 = ns2;
    out.s_off = dalloc<int>(ntiles + 1);
    out.s_ngt = dalloc<int>(ntiles + 1);
    out.s2 = dalloc<int>(ns2);
    k_tile_bounds<<<nblk(ntiles + 1), 256, 0, st>>>(ntiles, k2, ns2, 0, out.s_off);
    k_tile_bounds<<<nblk(ntiles + 1), 256, 0, st>>>(ntiles, k2, ns2, CLS_REST, out.s_ngt);
    if (ns2) k_key_ids<<<nblk(ns2), 256, 0, st>>>(ns2, k2, out.s2);
    HIPCHK(hipGetLastError());
    std::vector<int> h1o(ntiles + 1), so(ntiles + 1), sg(ntiles + 1);
    HIPCHK(hipMemcpyAsync(h1o.data(), out.h1_off, sizeof(int) * (ntiles + 1), hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(so.data(), out.s_off, sizeof(int) * (ntiles + 1), hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(sg.data(), out.s_ngt, sizeof(int) * (ntiles + 1), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    (void)hipFree(k2);
    nh.assign(ntiles, 0);
    ns.assign(ntiles, 0);
    ng.assign(ntiles, 0);
    for (int t = 0; t < ntiles; t++) {
        ng[t] = sg[t] - so[t];  // ghosts of the tile's cells: [s_off, first class-1 entry)
        nh[t] = h1o[t + 1] - h1o[t];
        ns[t] = so[t + 1] - so[t];
    }
    HIPCHK(hipMemcpyAsync(out.s_ngt, ng.data(), sizeof(int) * ntiles, hipMemcpyHostToDevice, st));
    HIPCHK(hipStreamSynchronize(st));
}

// Local indices of the neighbours of tile t's cells (T, then its H1 entries):
// one block per tile, the tile's H1 and S2 lists staged in LDS and searched.
__global__ void k_local_nb(int Nint, long long P, int nslot, const int* __restrict__ cnb,
                           const int* __restrict__ cside,
                           const int* __restrict__ tile_start, const int* __restrict__ h1_off,
                           const int* __restrict__ h1, const int* __restrict__ s_off, const int* __restrict__ s_ngt,
                           const int* __restrict__ s2, long long nh1tot, short* __restrict__ lt,
                           short* __restrict__ lh, unsigned long long* __restrict__ unresolved) {
    extern __shared__ int lists[];
    const int t = blockIdx.x;
    TileLists L;
    L.tb = tile_start[t];
    L.tc = tile_start[t + 1] - L.tb;
    const int h0 = h1_off[t], s0 = s_off[t];
    L.nh1 = h1_off[t + 1] - h0;
    L.ns2 = s_off[t + 1] - s0;
    L.ngt = s_ngt[t];
    L.Nint = Nint;
    int* sH1 = lists;
    int* sS2 = lists + L.nh1;
    for (int q = threadIdx.x; q < L.nh1; q += blockDim.x) sH1[q] = h1[h0 + q];
    for (int q = threadIdx.x; q < L.ns2; q += blockDim.x) sS2[q] = s2[s0 + q];
    L.sH1 = sH1;
    L.sS2 = sS2;
    __syncthreads();
    unsigned long long miss = 0;
    for (int q = threadIdx.x; q < L.tc + L.nh1; q += blockDim.x) {
        const int c = q < L.tc ? L.tb + q : sH1[q - L.tc];
        for (int k = 0; k < nslot; k++) {
            const int x = cnb[k * P + c];
            int l = -1;
            if (x != SLOT_EMPTY) {
                const int nb = x & SLOT_NB_MASK;
                l = local_of(L, nb);
                if (l < 0 || l > LT_LOCAL_MASK) {
                    miss++;
                    l = -1;
                } else {
                    l |= (int)((unsigned)x >> SLOT_TWIN_SHIFT) << LT_TWIN_SHIFT;
                    if (nb >= Nint) l |= LT_GHOST;
                    if ((cside[c] >> k) & 1) l |= LT_SIDE_R;
                }
            }
            if (q < L.tc) lt[k * P + c] = (short)l;
            else lh[k * nh1tot + h0 + (q - L.tc)] = (short)l;
        }
    }
    if (miss) atomicAdd(unresolved, miss);
}

// The 16-bit local tables of the final tiling (packed entries, LT_*: a tiling
// whose tiles stage more than 2^10 cells gets none, and the single-kernel step
// is unavailable for it; the LDS budgets keep every tile below that: at most
// about 900 staged cells in the mixed path, fewer in FP64).
void build_local_tables(int Nint, long long P, int nslot, const int* cnb, const int* cside, HaloTables& out,
                        hipStream_t st) {
    out.unresolved = 0;
    if (out.ntiles == 0 || out.smax > LT_LOCAL_MASK + 1) return;
    const size_t lds = sizeof(int) * ((size_t)out.h1max + out.s2max);
    if (lds > 64 * 1024) return;
    out.lt = dalloc<short>((size_t)nslot * P);
    out.lh = dalloc<short>((size_t)nslot * std::max(out.nh1, 1LL));
    HIPCHK(hipMemsetAsync(out.lt, 0xff, sizeof(short) * (size_t)nslot * P, st));
    unsigned long long* d_miss = dalloc<unsigned long long>(1);
    HIPCHK(hipMemsetAsync(d_miss, 0, sizeof(unsigned long long), st));
    k_local_nb<<<out.ntiles, 256, lds, st>>>(Nint, P, nslot, cnb, cside, out.tile_start, out.h1_off, out.h1, out.s_off,
                                             out.s_ngt, out.s2, std::max(out.nh1, 1LL), out.lt, out.lh, d_miss);
    HIPCHK(hipGetLastError());
    unsigned long long miss = 0;
    HIPCHK(hipMemcpyAsync(&miss, d_miss, sizeof(miss), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    (void)hipFree(d_miss);
    out.unresolved = (long long)miss;
}

}  // namespace

void build_halo_tables(int Nint, int NcPad, int nslot, const int* cnb, const int* cside, int TC, bool mixed,
                       size_t lds_budget,
                       HaloTables& out, hipStream_t st) {
    free_halo_tables(out);
    out.TC = TC;
    std::vector<int> ts;
    for (int c = 0; c < Nint; c += TC) ts.push_back(c);
    ts.push_back(Nint);
    std::vector<int> nh, ns, ng;
    for (int round = 0;; round++) {
        build_lists(Nint, NcPad, nslot, cnb, ts, out, nh, ns, ng, st);
        out.smax = out.fsmax = out.h1max = out.s2max = 0;
        const int ntiles = out.ntiles;
        std::vector<size_t> bytes(ntiles);
        for (int t = 0; t < ntiles; t++) {
            const int tc = ts[t + 1] - ts[t];
            out.smax = std::max(out.smax, tc + nh[t] + ns[t]);
            out.fsmax = std::max(out.fsmax, tc + nh[t] + ng[t]);
            out.h1max = std::max(out.h1max, nh[t]);
            out.s2max = std::max(out.s2max, ns[t]);
            bytes[t] = fused_lds_bytes(mixed, nslot, tc + nh[t] + ns[t], tc + nh[t] + ng[t]);
        }
        const size_t lds = fused_lds_bytes(mixed, nslot, out.smax, out.fsmax);
        if (lds <= lds_budget || round == 5) break;
        // halve the tiles whose own footprint is over budget (tightened a little
        // each round, since the block's LDS takes the maximum of each list)
        const size_t limit = lds_budget - lds_budget * (size_t)round / 20;
        std::vector<int> nt;
        nt.reserve(ts.size() + 64);
        int split = 0;
        for (int t = 0; t < ntiles; t++) {
            nt.push_back(ts[t]);
            const int tc = ts[t + 1] - ts[t];
            if (bytes[t] > limit && tc >= 32) {
                nt.push_back(ts[t] + tc / 2);
                split++;
            }
        }
        nt.push_back(Nint);
        if (!split) break;
        ts.swap(nt);
        free_halo_tables(out);
        out.TC = TC;
    }
    out.split_tiles = out.ntiles - (Nint + TC - 1) / TC;
    build_local_tables(Nint, NcPad, nslot, cnb, cside, out, st);
}

namespace {

// Entry e of the ring copy: the geometry of the H1 cell h1[e] at entry e of
// the same 64-cell block layout (geo_index), its 1/V at e.
template <typename T>
__global__ void k_ring_inputs(int nslot, long long nh1, const int* __restrict__ h1, const T* __restrict__ geo,
                              const T* __restrict__ iv, T* __restrict__ hgeo, T* __restrict__ hiv) {
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nh1) return;
    const long long g = h1[e];
    for (int p = 0; p < GEO_PLANES; p++)
        for (int k = 0; k < nslot; k++) hgeo[geo_index(nslot, p, k, e)] = geo[geo_index(nslot, p, k, g)];
    hiv[e] = iv[g];
}

}  // namespace

namespace {

// Tile holding cell c: the last tile whose first cell is <= c.
__device__ int tile_of(const int* __restrict__ tile_start, int ntiles, int c) {
    int lo = 0, hi = ntiles - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (tile_start[mid] <= c) lo = mid; else hi = mid - 1;
    }
    return lo;
}

__global__ void k_inst_tiles(const int* __restrict__ inst_off, int ninst, const int* __restrict__ tile_start,
                             int ntiles, int2* __restrict__ out) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= ninst) return;
    const int c0 = inst_off[b], c1 = inst_off[b + 1];
    out[b] = c1 > c0 ? make_int2(tile_of(tile_start, ntiles, c0), tile_of(tile_start, ntiles, c1 - 1))
                     : make_int2(0, -1);
}

}  // namespace

void build_instance_tiles(const int* inst_off, int ninst, HaloTables& t, hipStream_t st) {
    if (t.inst_tiles) (void)hipFree(t.inst_tiles);
    t.inst_tiles = dalloc<int2>((size_t)std::max(ninst, 1));
    if (ninst > 0 && t.ntiles > 0)
        k_inst_tiles<<<nblk(ninst), 256, 0, st>>>(inst_off, ninst, t.tile_start, t.ntiles, t.inst_tiles);
    HIPCHK(hipGetLastError());
}

void build_ring_inputs(int nslot, long long P, const double* cgeo, const double* iv, const float* cgeo32,
                       const float* iv32, HaloTables& t, hipStream_t st) {
    (void)P;
    const long long nh1 = (std::max(t.nh1, 1LL) + 63) / 64 * 64;  // whole 64-entry blocks
    const size_t ngeo = (size_t)GEO_PLANES * nslot * nh1;
    if (cgeo) {
        t.hgeo = dalloc<double>(ngeo);
        t.hiv = dalloc<double>((size_t)nh1);
        if (t.nh1 > 0)
            k_ring_inputs<double><<<nblk(t.nh1), 256, 0, st>>>(nslot, t.nh1, t.h1, cgeo, iv, t.hgeo, t.hiv);
    } else {
        t.hgeo32 = dalloc<float>(ngeo);
        t.hiv32 = dalloc<float>((size_t)nh1);
        if (t.nh1 > 0)
            k_ring_inputs<float><<<nblk(t.nh1), 256, 0, st>>>(nslot, t.nh1, t.h1, cgeo32, iv32, t.hgeo32, t.hiv32);
    }
    HIPCHK(hipGetLastError());
}

void free_halo_tables(HaloTables& t) {
    for (void* p : {(void*)t.tile_start, (void*)t.h1_off, (void*)t.h1, (void*)t.s_off, (void*)t.s_ngt, (void*)t.s2,
                    (void*)t.lt, (void*)t.lh, (void*)t.hgeo, (void*)t.hiv, (void*)t.hgeo32, (void*)t.hiv32,
                    (void*)t.inst_tiles})
        if (p) (void)hipFree(p);
    t = HaloTables{};
}

}  // namespace hcfd
