// passes_fp64.hip -- FP64 variants of the fused passes, the CFL reduction and
// the variant dispatch (see passes_impl.hpp for the method).
#include "dispatch.hpp"
#include "passes_impl.hpp"

namespace hcfd {

hipError_t launch_pass_a_mixed(const PassVariant& v, const PassArgs& a, int grid, int block, hipStream_t st);
hipError_t launch_pass_b_mixed(const PassVariant& v, const PassArgs& a, int grid, int block, hipStream_t st);

int pass_degree_bucket(int maxdeg) { return degree_bucket(maxdeg); }

bool pass_variant_supported(const PassVariant& v) {
    if (v.width != 1 && v.width != 8 && v.width != 64) return false;
    if (v.maxdeg != 3 && v.maxdeg != 4 && v.maxdeg != 8) return false;
    return v.precision == HCFD_FP64 || v.precision == HCFD_MIXED;
}

hipError_t launch_pass_a(const PassVariant& v, const PassArgs& a, int grid, int block, hipStream_t st) {
    if (!pass_variant_supported(v) || block > 256 || block % WAVE) return hipErrorInvalidValue;
    if (v.precision == HCFD_MIXED) return launch_pass_a_mixed(v, a, grid, block, st);
    return with_width(v.width, [&](auto W) {
        return with_degree(v.maxdeg, [&](auto MD) {
            return with_bool(v.cfl, [&](auto CF) {
                return with_bool(a.ninst > 1, [&](auto MU) {
                    return with_lanes(v.lpc, [&](auto LP) {
                        hipLaunchKernelGGL((k_pass_a<double, decltype(W)::value, decltype(MD)::value,
                                                     decltype(CF)::value, decltype(MU)::value, decltype(LP)::value>),
                                           dim3(grid), dim3(block), 0, st, a);
                        return hipGetLastError();
                    });
                });
            });
        });
    });
}

hipError_t launch_pass_b(const PassVariant& v, const PassArgs& a, int grid, int block, hipStream_t st) {
    if (!pass_variant_supported(v) || block > 256 || block % WAVE) return hipErrorInvalidValue;
    if (v.precision == HCFD_MIXED) return launch_pass_b_mixed(v, a, grid, block, st);
    return with_width(v.width, [&](auto W) {
        return with_degree(v.maxdeg, [&](auto MD) {
            return with_bool(a.ninst > 1, [&](auto MU) {
                hipLaunchKernelGGL((k_pass_b<false, decltype(W)::value, decltype(MD)::value, decltype(MU)::value>),
                                   dim3(grid), dim3(block), 0, st, a);
                return hipGetLastError();
            });
        });
    });
}

// Per-instance CFL reduction over the tile partials (maximum rates, see cfl_rate;
// block b <-> instance b): dt = cfl * (1 / max rate), the value cfl * min_i(1/rate_i).
__global__ __launch_bounds__(1024) void k_dt_reduce(PassArgs a, double cfl, double* __restrict__ dt) {
    __shared__ double sh[1024 / WAVE];
    const int b = blockIdx.x;
    const int c0 = a.inst_off[b], c1 = a.inst_off[b + 1];
    double m = 0.0;
    if (c1 > c0) {
        // tiles of tile_cells cells, or the variable tiles of tile_start (one
        // instance: every tile; several: the range built with the tiles, else
        // thread 0 searches and shares it)
        __shared__ int s_t[2];
        int t0, t1;
        if (!a.tile_start) {
            t0 = c0 / a.tile_cells;
            t1 = (c1 - 1) / a.tile_cells;
        } else if (a.ninst == 1) {
            t0 = 0;
            t1 = a.ntiles - 1;
        } else if (a.inst_tiles) {
            const int2 r = a.inst_tiles[b];
            t0 = r.x;
            t1 = r.y;
        } else {
            if (threadIdx.x == 0) {
                for (int e = 0; e < 2; e++) {
                    const int c = e == 0 ? c0 : c1 - 1;
                    int lo = 0, hi = a.ntiles - 1;
                    while (lo < hi) {
                        const int mid = (lo + hi + 1) >> 1;
                        if (a.tile_start[mid] <= c) lo = mid; else hi = mid - 1;
                    }
                    s_t[e] = lo;
                }
            }
            __syncthreads();
            t0 = s_t[0];
            t1 = s_t[1];
        }
        if (a.ninst == 1) {
            // one instance: every partial is its tile's first; four loads in flight per lane
            double m1 = 0.0, m2 = 0.0, m3 = 0.0;
            int tile = t0 + (int)threadIdx.x;
            const int s = blockDim.x;
            for (; tile + 3 * s <= t1; tile += 4 * s) {
                m = fmax(m, a.pmin_first[tile]);
                m1 = fmax(m1, a.pmin_first[tile + s]);
                m2 = fmax(m2, a.pmin_first[tile + 2 * s]);
                m3 = fmax(m3, a.pmin_first[tile + 3 * s]);
            }
            for (; tile <= t1; tile += s) m = fmax(m, a.pmin_first[tile]);
            m = fmax(fmax(m, m1), fmax(m2, m3));
        } else {
            // tiles strictly inside the range start in this instance (their first
            // partial); the range's two end tiles may start or end in another
            for (int tile = t0 + (int)threadIdx.x; tile <= t1; tile += blockDim.x) {
                if (tile > t0 && tile < t1) {
                    m = fmax(m, a.pmin_first[tile]);
                    continue;
                }
                const int f0 = a.tile_start ? a.tile_start[tile] : tile * a.tile_cells;
                const int l0 = a.tile_start ? a.tile_start[tile + 1] - 1 : min(f0 + a.tile_cells, a.Nint) - 1;
                const int bf = instance_from(a.inst_off, a.chunk_inst, f0);
                const int bl = instance_from(a.inst_off, a.chunk_inst, l0);
                if (bf == b) m = fmax(m, a.pmin_first[tile]);
                else if (bl == b) m = fmax(m, a.pmin_last[tile]);
            }
        }
    }
    m = block_max(m, sh);
    if (threadIdx.x == 0) {
        m = fmax(m, __longlong_as_double((long long)a.dtacc[b]));
        dt[b] = cfl * (1.0 / m);
        a.dtacc[b] = 0ull;
    }
}

// One instance: the tile partials split over the grid's blocks; each folds its
// maximum into dtacc[0] (device-scope atomic on the rate's bit pattern, which
// orders like the value), and the last block to arrive (ticket dtacc[1]) forms
// dt = cfl * (1 / max) and resets both slots for the next step.
__global__ __launch_bounds__(256) void k_dt_reduce_one(PassArgs a, double cfl, double* __restrict__ dt) {
    __shared__ double sh[256 / WAVE];
    const int G = gridDim.x;
    const int per = (a.ntiles + G - 1) / G;
    const int t0 = blockIdx.x * per, t1 = min(t0 + per, a.ntiles);
    double m = 0.0, m1 = 0.0;
    const int s = blockDim.x;
    int tile = t0 + (int)threadIdx.x;
    for (; tile + s < t1; tile += 2 * s) {
        m = fmax(m, a.pmin_first[tile]);
        m1 = fmax(m1, a.pmin_first[tile + s]);
    }
    if (tile < t1) m = fmax(m, a.pmin_first[tile]);
    m = block_max(fmax(m, m1), sh);
    if (threadIdx.x == 0) {
        atomicMax(&a.dtacc[0], dbits(m));
        __threadfence();
        if (atomicAdd(&a.dtacc[1], 1ull) == (unsigned long long)(G - 1)) {
            __threadfence();
            const double mx = __longlong_as_double((long long)atomicAdd(&a.dtacc[0], 0ull));
            dt[0] = cfl * (1.0 / mx);
            atomicExch(&a.dtacc[0], 0ull);
            atomicExch(&a.dtacc[1], 0ull);
        }
    }
}

hipError_t launch_dt_reduce(const PassArgs& a, double cfl, double* dt, hipStream_t st) {
    if (a.ninst == 1) {
        // about 1,024 partials per block, at most 64 blocks
        const int G = std::max(1, std::min(64, a.ntiles / 1024));
        hipLaunchKernelGGL(k_dt_reduce_one, dim3(G), dim3(256), 0, st, a, cfl, dt);
    } else {
        hipLaunchKernelGGL(k_dt_reduce, dim3(a.ninst), dim3(1024), 0, st, a, cfl, dt);
    }
    return hipGetLastError();
}

}  // namespace hcfd
