// kernels.hip - the per-tick physics kernels for gfx950.
//
// Every kernel walks its kernel class's blocks over the batch (grid-stride), so
// one launch serves one scene or many. Per tick (Simulation::step,
// crates/embloks-core/src/simulation.rs:63-72, in the order of
// GpuSim::encode_ticks, crates/embloks-gpu/src/sim.rs:607-661):
//
//   trim        battery gain feedback (battery.rs:260-281): trim_block's
//               blocks in the tiled push's launch, then k_trim_gain
//   k_push      electron substep 0: gather, impressed field, P3M close force,
//               implicit drag, step clamp, confinement (physics.rs:124-162)
//   k_rest      electron substeps 1.. when the tick needs more than one
//   k_count, k_scan, k_scatter, k_canon
//               counting sort of the electrons by cell (close_range.rs:258-286)
//   k_deposit_smooth
//               integer CIC deposit of rho and J plus the 1-2-1 smoothing
//               (fields.rs:18-63, maxwell.rs:280-316), cell-binned current
//               (physics.rs:34-43) and the speed maximum for the next tick
//   k_fdtd      leapfrog + absorbing margin (maxwell.rs:235-270), then the
//               right-hand sides of the Gauss solve
//   dst GEMMs   (gemm.hip) and k_correct: E -= grad(phi) (maxwell.rs:176-204)

#include "device.hpp"
#include "gemm.hpp"

namespace emb {

// The sorted snapshot holds (x, y, vx, vy) per electron in one 16-byte entry:
// the deposit reads an electron with one load, a pair loop its position half.
__device__ __forceinline__ float2 snap_pos(const DevScene &S, int j) {
    const float4 e = S.snap[j];
    return make_float2(e.x, e.y);
}
__device__ __forceinline__ float2 snap_vel(const DevScene &S, int j) {
    const float4 e = S.snap[j];
    return make_float2(e.z, e.w);
}

// ---------------------------------------------------------------------------------------
// Counting sort by clamped cell (close_range.rs:258-286; particles.wgsl:23-85).
// Ranks within a cell come from atomics; k_canon then orders each cell's slots by
// electron index, so the sorted order (and every sum over it) is deterministic.

__global__ __launch_bounds__(256) void k_count(Batch b) {
    const int lane = threadIdx.x & 63;
    for (int blk = blockIdx.x; blk < b.total[KC_ELECTRONS]; blk += gridDim.x) {
        const int s = scene_of(b, KC_ELECTRONS, blk);
        const DevScene &S = b.scenes[s];
        const int i = block_in_scene(b, KC_ELECTRONS, s, blk) * 256 + (int)threadIdx.x;
        const bool valid = i < S.n;   // a prefix of the wave's lanes
        const int cell = valid ? cell_of(S, S.cxy[i]) : -1;
        // One atomic per run of consecutive lanes in the same cell (electrons mostly
        // arrive grouped by cell), all runs of the wave at once.
        const int prev = __shfl_up(cell, 1, 64);
        const unsigned long long heads = __ballot(valid && (lane == 0 || prev != cell));
        const int n_valid = __popcll(__ballot(valid));
        const unsigned long long upto = (2ull << lane) - 1ull;   // lanes 0..lane
        const int head = 63 - __clzll(heads & upto);
        int base = 0;
        if (valid && head == lane) {
            const unsigned long long later = heads & ~upto;
            const int next = later ? __ffsll((long long)later) - 1 : n_valid;
            base = atomicAdd(&S.count[cell], next - lane);
        }
        base = __shfl(base, valid ? head : 0, 64);
        const int rank = base + (lane - head);
        if (valid) {
            S.erank[i] = rank;
            S.ecell[i] = cell;
        }
    }
}

// Exclusive scan of 256 values, one per thread; returns the exclusive prefix.
__device__ __forceinline__ int block_scan_256(int v, int *wsum, int &total) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int inc = v;
    for (int o = 1; o < 64; o <<= 1) {
        const int t = __shfl_up(inc, o, 64);
        if (lane >= o) inc += t;
    }
    if (lane == 63) wsum[w] = inc;
    __syncthreads();
    int off = 0;
    total = 0;
    for (int k = 0; k < 4; k++) {
        const int ws = wsum[k];
        if (k < w) off += ws;
        total += ws;
    }
    __syncthreads();
    return off + inc - v;
}

// Inclusive prefix sum over the 64 lanes of a wave.
__device__ __forceinline__ int wave_scan_incl(int v) {
    const int lane = threadIdx.x & 63;
    for (int o = 1; o < 64; o <<= 1) {
        const int u = __shfl_up(v, o, 64);
        if (lane >= o) v += u;
    }
    return v;
}

// Block-wide exclusive scan of one value per thread over THREADS threads
// (every thread of the block calls it); returns the thread's prefix and sets
// `total`.
template <int THREADS>
__device__ __forceinline__ int block_scan(int v, int *wsum, int &total) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int incl = wave_scan_incl(v);
    if (lane == 63) wsum[w] = incl;
    __syncthreads();
    int base = 0, all = 0;
    for (int q = 0; q < THREADS / 64; q++) {
        if (q < w) base += wsum[q];
        all += wsum[q];
    }
    __syncthreads();
    total = all;
    return base + incl - v;
}

__global__ __launch_bounds__(256) void k_scan(Batch b) {
    __shared__ int wsum[4];
    for (int blk = blockIdx.x; blk < b.total[KC_SCAN]; blk += gridDim.x) {
        const int s = scene_of(b, KC_SCAN, blk);
        const DevScene &S = b.scenes[s];
        const int tile = block_in_scene(b, KC_SCAN, s, blk);
        const int cells = S.cols * S.rows;
        const int base = tile * SCAN_TILE + (int)threadIdx.x * 4;
        int v[4], sum = 0;
        for (int q = 0; q < 4; q++) {
            v[q] = base + q < cells ? S.count[base + q] : 0;
            sum += v[q];
        }
        int total;
        int run = block_scan_256(sum, wsum, total);
        for (int q = 0; q < 4; q++) {
            if (base + q < cells) {
                S.lstart[base + q] = run;
                S.count[base + q] = 0;
            }
            run += v[q];
        }
        if (threadIdx.x == 0) S.toff[tile] = total;
        // This scan tile's share of the push's dense-tile flags, cleared for
        // k_canon to set, four flags per store (the flag array's allocation is
        // padded to 256 bytes, so its last word is whole).
        const int ntiles = (cells + SCAN_TILE - 1) / SCAN_TILE;
        const int ptiles = ((S.cols + PUSH_TILE - 1) / PUSH_TILE) * ((S.rows + PUSH_TILE - 1) / PUSH_TILE);
        const int words = (ptiles + 3) / 4, per = (words + ntiles - 1) / ntiles;
        uint32_t *flags = reinterpret_cast<uint32_t *>(S.tile_dense);
        for (int q = tile * per + (int)threadIdx.x; q < min(words, (tile + 1) * per); q += 256) flags[q] = 0u;
        __syncthreads();
    }
}

// One block per scene: the scan tiles' totals become their offsets. The totals
// are read and written back coalesced through LDS, TOP_CHUNK at a time, and each
// of the 1024 threads scans TOP_RUN consecutive ones of a chunk (a 4096^2 grid's
// 16k tiles take two chunks; a 512^2 grid's 256 tiles one, on 32 threads).
constexpr int TOP_THREADS = 1024, TOP_RUN = 8, TOP_CHUNK = TOP_THREADS * TOP_RUN;
__global__ __launch_bounds__(TOP_THREADS) void k_scan_top(Batch b) {
    __shared__ int buf[TOP_CHUNK];
    __shared__ int wsum[TOP_THREADS / 64];
    for (int blk = blockIdx.x; blk < b.total[KC_SCENE]; blk += gridDim.x) {
        const int s = scene_of(b, KC_SCENE, blk);
        const DevScene &S = b.scenes[s];
        const int ntiles = (S.cols * S.rows + SCAN_TILE - 1) / SCAN_TILE;
        int carry = 0;
        for (int c0 = 0; c0 < ntiles; c0 += TOP_CHUNK) {
            const int m = min(TOP_CHUNK, ntiles - c0);
            for (int i = threadIdx.x; i < m; i += TOP_THREADS) buf[i] = S.toff[c0 + i];
            __syncthreads();
            const int k0 = (int)threadIdx.x * TOP_RUN;
            int v[TOP_RUN], sum = 0;
#pragma unroll
            for (int q = 0; q < TOP_RUN; q++) {
                v[q] = k0 + q < m ? buf[k0 + q] : 0;
                sum += v[q];
            }
            int chunk_total;
            int run = carry + block_scan<TOP_THREADS>(sum, wsum, chunk_total);
#pragma unroll
            for (int q = 0; q < TOP_RUN; q++) {
                if (k0 + q < m) buf[k0 + q] = run;
                run += v[q];
            }
            __syncthreads();
            for (int i = threadIdx.x; i < m; i += TOP_THREADS) S.toff_pre[c0 + i] = buf[i];
            carry += chunk_total;
            __syncthreads();
        }
        if (threadIdx.x == 0) {
            S.ctr->heavy_n = 0;
            S.ctr->dense_n = 0;
        }
        __syncthreads();
    }
}

// Each electron to its slot: its cell's start plus its rank there. With TOP
// (every scene of the batch has at most 256 scan tiles), each block first scans
// the tile totals itself, in place of k_scan_top's launch, and the scene's
// first block stores the offsets for k_canon and clears the list counters.
template <bool TOP>
__global__ __launch_bounds__(256) void k_scatter(Batch b) {
    __shared__ int ltoff[256];
    __shared__ int wsum[4];
    for (int blk = blockIdx.x; blk < b.total[KC_ELECTRONS]; blk += gridDim.x) {
        const int s = scene_of(b, KC_ELECTRONS, blk);
        const DevScene &S = b.scenes[s];
        const int local = block_in_scene(b, KC_ELECTRONS, s, blk);
        const int i = local * 256 + (int)threadIdx.x;
        if (TOP) {
            const int cells = S.cols * S.rows, ntiles = (cells + SCAN_TILE - 1) / SCAN_TILE;
            const int v = (int)threadIdx.x < ntiles ? S.toff[threadIdx.x] : 0;
            int total;
            const int pre = block_scan_256(v, wsum, total);
            ltoff[threadIdx.x] = pre;
            if (local == 0) {
                if ((int)threadIdx.x < ntiles) S.toff_pre[threadIdx.x] = pre;
                if (threadIdx.x == 0) {
                    S.ctr->heavy_n = 0;
                    S.ctr->dense_n = 0;
                }
            }
            __syncthreads();
            if (i < S.n) {
                const int cell = S.ecell[i];
                S.order[S.lstart[cell] + ltoff[cell / SCAN_TILE] + S.erank[i]] = i;
            }
            __syncthreads();   // the next block's scan reuses ltoff
        } else if (i < S.n) {
            S.order[scan_start(S, S.ecell[i]) + S.erank[i]] = i;
        }
    }
}

// Per cell: absolute start, electron-index order within the cell, and the sorted
// snapshot of positions and velocities.
__device__ __forceinline__ void canon_cell(const DevScene &S, int c, int s0, int s1) {
    int *order = S.order;
    for (int a = s0 + 1; a < s1; a++) {
        const int key = order[a];
        int p = a - 1;
        while (p >= s0 && order[p] > key) {
            order[p + 1] = order[p];
            p--;
        }
        order[p + 1] = key;
    }
    for (int a = s0; a < s1; a++) {
        const int e = order[a];
        S.scell[a] = c;
        S.snap[a] = make_float4(S.fx[e], S.fy[e], S.vx[e], S.vy[e]);
    }
}

// 256 cells per block: their slots are loaded into LDS, each cell's slots are
// ordered by electron index there, and the ordered slots and the snapshot are
// written back with coalesced stores. Blocks whose cells hold more than
// CANON_CAP electrons sort in global memory instead. 4096 (256 cells at 16
// electrons, 32 KB of LDS) measured 29 us on the 512^2 scene against 33 us at
// 6144 (more blocks resident) and 134 us at 2048 (most blocks fall back).
constexpr int CANON_CAP = 4096;

__global__ __launch_bounds__(256) void k_canon(Batch b) {
    __shared__ int lorder[CANON_CAP], lsorted[CANON_CAP];
    __shared__ int lstart[257];
    __shared__ uint8_t lcell[CANON_CAP];   // block-local cell of each slot
    __shared__ int wnew[5];                // new dense tiles per wave, then the block's list base
    for (int blk = blockIdx.x; blk < b.total[KC_CELLS]; blk += gridDim.x) {
        const int s = scene_of(b, KC_CELLS, blk);
        const DevScene &S = b.scenes[s];
        const int cells = S.cols * S.rows;
        const int c0 = block_in_scene(b, KC_CELLS, s, blk) * 256;
        const int c1 = min(c0 + 256, cells);
        const int S0 = scan_start(S, c0), S1 = scan_start(S, c1);
        const int c = c0 + (int)threadIdx.x;
        const bool mine = c < c1;
        const int s0 = mine ? scan_start(S, c) : 0, s1 = mine ? scan_start(S, c + 1) : 0;
        const bool in_lds = S1 - S0 <= CANON_CAP;
        // The block's slots into LDS first: their loads are in flight during the
        // dense-tile atomics below, whose results each wait for.
        if (in_lds) {
            for (int q = S0 + (int)threadIdx.x; q < S1; q += 256) lorder[q - S0] = S.order[q];
            lstart[threadIdx.x] = (mine ? s0 : S1) - S0;
            if (threadIdx.x == 0) lstart[256] = S1 - S0;
        }
        // A push tile with a crowded cell runs the cell mode (k_push_tiled);
        // the others go electron by electron (k_push_sparse). The first cell to
        // flag its tile lists it (the flag's byte set in its word by an atomic
        // OR; k_scan cleared the words). Consecutive lanes hold consecutive
        // rows of a column, so only the first crowded lane of each run of one
        // tile sets the flag: one atomic per tile column, not per cell.
        {
            int tile = -1;
            if (mine && s1 - s0 >= CELL_MODE_MIN) {
                const int cc = c / S.rows, rr = c - cc * S.rows;
                tile = (cc / PUSH_TILE) * ((S.rows + PUSH_TILE - 1) / PUSH_TILE) + rr / PUSH_TILE;
            }
            const int lane = threadIdx.x & 63;
            const int prev = __shfl_up(tile, 1, 64);
            bool first = false;
            if (tile >= 0 && (lane == 0 || prev != tile)) {
                const unsigned bit = 1u << (8 * (tile & 3));
                first = (atomicOr(reinterpret_cast<unsigned *>(S.tile_dense + (tile & ~3)), bit) & bit) == 0;
            }
            // the block's new tiles take consecutive list places: one counter
            // atomic per block (atomics on one word queue at its L2 channel)
            const unsigned long long m = __ballot(first);
            const int w = threadIdx.x >> 6;
            if (lane == 0) wnew[w] = __popcll(m);
            __syncthreads();
            if (threadIdx.x == 0) {
                int sum = 0;
                for (int q = 0; q < 4; q++) {
                    const int c4 = wnew[q];
                    wnew[q] = sum;
                    sum += c4;
                }
                wnew[4] = sum ? atomicAdd(&S.ctr->dense_n, sum) : 0;
            }
            __syncthreads();
            if (first) S.dense_tiles[wnew[4] + wnew[w] + __popcll(m & ((1ull << lane) - 1ull))] = tile;
        }
        if (mine) {
            S.cstart[c] = s0;
            if (c == cells - 1) S.cstart[cells] = S.n;
            // Crowded cells get their own blocks (k_push_heavy); the list's order
            // does not matter, each cell's result depends on the cell alone.
            if (s1 - s0 >= HEAVY_CELL) S.heavy_cells[atomicAdd(&S.ctr->heavy_n, 1)] = c;
        }
        if (!in_lds) {
            if (mine) canon_cell(S, c, s0, s1);
            __syncthreads();
            continue;
        }
        // each cell marks its slots with its block-local index (a slot's cell
        // without a search over the starts)
        if (mine)
            for (int a = s0 - S0; a < s1 - S0; a++) lcell[a] = (uint8_t)threadIdx.x;
        __syncthreads();
        // Rank sort within each cell, one thread per slot: an electron's place is
        // the number of smaller indices in its cell (the indices are distinct).
        // A crowded cell costs each of its lanes one pass over the cell, where a
        // thread's insertion sort would cost it the square.
        for (int a = threadIdx.x; a < S1 - S0; a += 256) {
            const int ci = lcell[a];
            const int b0 = lstart[ci], b1 = lstart[ci + 1];
            const int key = lorder[a];
            int rank = 0;
            for (int j = b0; j < b1; j++) rank += lorder[j] < key ? 1 : 0;
            lsorted[b0 + rank] = key;
        }
        __syncthreads();
        // Four slots per thread per round: their gathers are all in flight at once.
        for (int q0 = S0 + (int)threadIdx.x; q0 < S1; q0 += 4 * 256) {
            int e[4];
            int cel[4];
            float fxv[4], fyv[4], vxv[4], vyv[4];
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const int q = q0 + u * 256;
                e[u] = q < S1 ? lsorted[q - S0] : 0;
            }
#pragma unroll
            for (int u = 0; u < 4; u++) {
                cel[u] = S.ecell[e[u]];
                fxv[u] = S.fx[e[u]];
                fyv[u] = S.fy[e[u]];
                vxv[u] = S.vx[e[u]];
                vyv[u] = S.vy[e[u]];
            }
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const int q = q0 + u * 256;
                if (q >= S1) continue;
                S.order[q] = e[u];
                S.scell[q] = cel[u];
                S.snap[q] = make_float4(fxv[u], fyv[u], vxv[u], vyv[u]);
            }
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------------------
// Close-range tables (close_range.rs:244-254, common.wgsl:88-100).

// Channel k of the square table at fractional indices, as table_bilinear
// (close_range.rs:245-254). The blend's products and sums are written as
// explicit FMAs, so every kernel that reads the table rounds it the same way,
// whichever products the compiler would otherwise fuse.
__device__ __forceinline__ float square_floor(const DevScene &S, int k, float u, float v) {
    const int n = S.square_n;
    const int a = clampi((int)floorf(u), 0, n - 2), b = clampi((int)floorf(v), 0, n - 2);
    const float fu = u - (float)a, fv = v - (float)b, gu = 1.0f - fu, gv = 1.0f - fv;
    const float *t = S.square + ((size_t)a * n + b) * 3 + k;
    const size_t n3 = (size_t)n * 3;
    float r = __fmul_rn(__fmul_rn(t[0], gu), gv);
    r = __fmaf_rn(__fmul_rn(t[n3], fu), gv, r);
    r = __fmaf_rn(__fmul_rn(t[3], gu), fv, r);
    return __fmaf_rn(__fmul_rn(t[n3 + 3], fu), fv, r);
}

// ---------------------------------------------------------------------------------------
// Battery trim (battery.rs:260-281; fields.wgsl:300-343). TRIM_BLOCKS blocks per
// stack. The potential a terminal reads is the node potential plus the
// close-range potential of the ions and electrons near it (Simulation::potential,
// simulation.rs:48-55; CloseRange::potential, close_range.rs:415-464).

// One block's share of a battery stack's terminal potentials (block `blk` of
// KC_STACKS: TRIM_BLOCKS blocks per stack, THREADS threads each), written to
// trim_part for k_trim_gain. The trim only reads the potential and the sorted
// snapshot of the tick before, so its blocks run in the tiled push's launch,
// beside the push's own blocks; the gain follows in a kernel of its own, before
// the kick reads it.
template <int THREADS>
__device__ __forceinline__ void trim_block(const Batch &b, int blk) {
    __shared__ float part[2][THREADS / 64];
    const int s = scene_of(b, KC_STACKS, blk);
    const DevScene &S = b.scenes[s];
    const int local = block_in_scene(b, KC_STACKS, s, blk);
    const int st = local / TRIM_BLOCKS, tb = local - st * TRIM_BLOCKS;
    const uint32_t t = S.ctr->tick;
    if (local == 0 && threadIdx.x == 0) S.ctr->vmax_bits[(t + 1) & 1] = 0;
    if (local == 0)   // k_rest's queue of this tick starts empty
        for (int q = threadIdx.x; q < rest_sync_len(S.max_substeps); q += THREADS) S.rest_sync[q] = 0;
    if (st >= S.n_stacks) return;
    const StackInfo I = S.stacks[st];
    if (!I.both_wired || I.amplitude == 0.0f) {
        if (tb == 0 && threadIdx.x == 0) {
            S.has_gain[st] = 0;
            S.gains[st] = 1.0f;
            S.stack_read[3 * st + 2] = 0.0f;   // the readings stay as they were (fields.wgsl:321-324)
        }
        return;
    }
    const int R = S.close_on ? S.reach : -1;
    const int side = 2 * R + 1;
    const int per = 1 + (S.close_on ? side * side : 0);
    const int lists = I.low_n + I.high_n;
    float acc_lo = 0.0f, acc_hi = 0.0f;
    // Work item = (terminal cell, neighbour cell, one of TRIM_SPLIT lanes over
    // that cell's electrons), so no thread walks a whole cell's electrons; the
    // stack's TRIM_BLOCKS blocks take every TRIM_BLOCKS-th stride of items.
    constexpr int TRIM_SPLIT = 4;
    for (int it = tb * THREADS + (int)threadIdx.x; it < lists * per * TRIM_SPLIT; it += THREADS * TRIM_BLOCKS) {
        const int part_i = it % TRIM_SPLIT, item = it / TRIM_SPLIT;
        const int li = item / per, sub = item - li * per;
        const bool high = li >= I.low_n;
        const int cell = S.stack_cells[high ? I.high_off + (li - I.low_n) : I.low_off + li];
        const int c = cell / S.rows, r = cell - c * S.rows;
        float v = 0.0f;
        if (sub == 0) {
            if (part_i == 0) {
                v = S.phi[(size_t)(c + S.margin) * S.pny + (r + S.margin)];
                if (S.close_on) v += S.ion_pot[cell];
            }
        } else {
            const int cc = c + (sub - 1) / side - R, rr = r + (sub - 1) % side - R;
            if (cc >= 0 && rr >= 0 && cc < S.cols && rr < S.rows) {
                const int nb = cc * S.rows + rr;
                // Offset of the electron from the centre of cell (c, r).
                const float ox = (float)(cc - c) - 0.5f + S.half, oy = (float)(rr - r) - 0.5f + S.half;
                for (int k = S.cstart[nb] + part_i; k < S.cstart[nb + 1]; k += TRIM_SPLIT) {
                    const float2 f = snap_pos(S, k);
                    const float u = (ox + f.x) * S.square_samples;
                    const float w = (oy + f.y) * S.square_samples;
                    v += square_floor(S, 0, u, w);
                }
                v *= S.k_q;
            }
        }
        if (high) acc_hi += v;
        else acc_lo += v;
    }
    for (int o = 32; o > 0; o >>= 1) {
        acc_lo += __shfl_xor(acc_lo, o, 64);
        acc_hi += __shfl_xor(acc_hi, o, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        part[0][threadIdx.x >> 6] = acc_lo;
        part[1][threadIdx.x >> 6] = acc_hi;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        float lo = 0.0f, hi = 0.0f;
        for (int w = 0; w < THREADS / 64; w++) {
            lo += part[0][w];
            hi += part[1][w];
        }
        float *mine = S.trim_part + 2 * ((size_t)st * TRIM_BLOCKS + tb);
        mine[0] = lo;
        mine[1] = hi;
    }
    __syncthreads();
}

// The stack's gain from its blocks' parts (trim_block, in the kernel before),
// added in block order, so the result does not depend on which block ended
// last. A kernel of its own rather than a last-block handshake: the
// release/acquire of a device-scope counter writes back and invalidates the L2
// of the block's XCD, once per trim block.
__global__ __launch_bounds__(64) void k_trim_gain(Batch b) {
    for (int blk = blockIdx.x; blk < b.total[KC_STACKS]; blk += gridDim.x) {
        const int s = scene_of(b, KC_STACKS, blk);
        const DevScene &S = b.scenes[s];
        const int local = block_in_scene(b, KC_STACKS, s, blk);
        const int st = local / TRIM_BLOCKS, tb = local - st * TRIM_BLOCKS;
        if (tb != 0 || st >= S.n_stacks || threadIdx.x != 0) continue;
        const StackInfo I = S.stacks[st];
        if (!I.both_wired || I.amplitude == 0.0f) continue;   // trim_block set the stack's readings
        const uint32_t t = S.ctr->tick;
        float lo = 0.0f, hi = 0.0f;
        const float *parts = S.trim_part + 2 * (size_t)st * TRIM_BLOCKS;
        for (int q = 0; q < TRIM_BLOCKS; q++) {
            lo += parts[2 * q];
            hi += parts[2 * q + 1];
        }
        const float v_low = lo / (float)I.low_n, v_high = hi / (float)I.high_n;
        const int cells = S.cols * S.rows;
        float isum = 0.0f;
        for (int k = 0; k < I.cur_n; k++) {
            const int cell = S.stack_cells[I.cur_off + k];
            const float ix = (float)S.cur[cell] * S.i_from_fixed;
            const float iy = (float)S.cur[cells + cell] * S.i_from_fixed;
            isum += ix * I.dir_x + iy * I.dir_y;
        }
        const float current = isum / (float)I.cur_n;
        const float target = S.target_tab[(size_t)(t % TICK_TABLE) * S.n_stacks + st];
        const float delivered = v_high - v_low + current * I.internal_resistance;
        float gain = S.has_gain[st] ? S.gains[st] : 1.0f;
        gain += S.trim_rate * (target - delivered) * target / (I.amplitude * I.amplitude);
        S.gains[st] = fminf(fmaxf(gain, S.gain_low), S.gain_high);
        S.has_gain[st] = 1;
        S.stack_read[3 * st] = v_low;
        S.stack_read[3 * st + 1] = v_high;
        S.stack_read[3 * st + 2] = 1.0f;
    }
}

// ---------------------------------------------------------------------------------------
// Electron push (physics.rs:124-162; particles.wgsl:193-227).

// Bilinear interpolation of a (w, h) array at array coordinates (i + fi, j + fj)
// with fi, fj in [0, 1], the index clamped like Maxwell::bilinear
// (maxwell.rs:381-390).
__device__ __forceinline__ float gather(const float *a, int w, int h, int i, float fi, int j, float fj) {
    const int ic = clampi(i, 0, w - 2), jc = clampi(j, 0, h - 2);
    const float fx = (float)(i - ic) + fi, fy = (float)(j - jc) + fj;
    const float *p = a + (size_t)ic * h + jc;
    return p[0] * (1.0f - fx) * (1.0f - fy) + p[h] * fx * (1.0f - fy) + p[1] * (1.0f - fx) * fy + p[h + 1] * fx * fy;
}

__device__ __forceinline__ bool conductive(const DevScene &S, int c, int r) {
    return (S.info[c * S.rows + r].flags & F_CONDUCTIVE) != 0u;
}

// physics.rs:83-115 for one axis, in coordinates relative to the start cell:
// p is the target's offset from the start cell's low face. A move that stays
// in the cell or crosses an open face is kept; a closed face (or the grid edge)
// reflects it with restitution, 1e-6 inside the cell (WALL_GAP, physics.rs:10).
__device__ __forceinline__ void resolve(const DevScene &S, int axis, int start, int other, float &p, float &v) {
    const int step = (int)floorf(p);
    if (step == 0) return;
    const bool forward = step > 0;
    const int face = forward ? start : start - 1;
    const int n = axis == 0 ? S.cols : S.rows, n_other = axis == 0 ? S.rows : S.cols;
    bool open = face >= 0 && face < n - 1 && other >= 0 && other < n_other;
    if (open) open = axis == 0 ? (conductive(S, face, other) && conductive(S, face + 1, other))
                               : (conductive(S, other, face) && conductive(S, other, face + 1));
    if (open) return;
    const float wall = forward ? 1.0f : 0.0f;
    const float bounced = wall - S.restitution * (p - wall);
    p = fminf(fmaxf(bounced, 1e-6f), 1.0f - 1e-6f);
    v *= -S.restitution;
}

// Split a cell-relative coordinate back into cell and fraction in [0, 1).
__device__ __forceinline__ void renormalise(int &cell, float &f) {
    const float fl = floorf(f);
    cell += (int)fl;
    f -= fl;
    if (f >= 1.0f) f = 0.99999994f;   // -tiny + 1 rounded up: stay below the face
}

__device__ __forceinline__ int substeps_of(const DevScene &S, uint32_t t) {
    const float vmax = __uint_as_float(S.ctr->vmax_bits[t & 1]);
    const float want = ceilf(vmax * S.dt / S.max_step);
    int sub = want >= 1.0f ? (int)fminf(want, (float)S.max_substeps) : 1;
    return sub < 1 ? 1 : sub;
}

// Positions in the close-range sums are 24-bit fixed point: cell offset times
// 2^24 plus the fraction rounded to 2^-24 of a cell. Pair offsets are then exact
// integer differences, so two electrons 0.01 cells apart keep their separation
// to ~1e-5 of itself whatever frame the positions are staged in.
__device__ __forceinline__ int fix24(float f) { return __float2int_rn(f * 16777216.0f); }
__device__ __forceinline__ int2 fixed_pos(int dc, int dr, float fx, float fy) {
    return make_int2(dc * (1 << 24) + fix24(fx), dr * (1 << 24) + fix24(fy));
}

// Lorentz force of the gathered fields plus the battery's impressed field, as an
// acceleration (physics.rs:49-51, 139-149; battery.rs:210-228).
__device__ __forceinline__ void field_accel(const DevScene &S, int hc, int hr, float px, float py, float vx, float vy,
                                            uint32_t flags, int k, float h, uint32_t t, float &ax, float &ay) {
    const int m = S.margin;
    // Maxwell::gather (maxwell.rs:336-350): Ex at (x + m, y - 1/2 + m), Ey at
    // (x - 1/2 + m, y + m), Bz at (x + m, y + m), each from its staggered grid.
    const int hx = px >= 0.5f ? hc : hc - 1, hy = py >= 0.5f ? hr : hr - 1;
    const float qx = px >= 0.5f ? px - 0.5f : px + 0.5f, qy = py >= 0.5f ? py - 0.5f : py + 0.5f;
    float ex = gather(S.ex, S.nx + 1, S.ny, hc + m, px, hy + m, qy);
    float ey = gather(S.ey, S.nx, S.ny + 1, hx + m, qx, hr + m, py);
    const float bz = gather(S.bz, S.nx + 1, S.ny + 1, hc + m, px, hr + m, py);
    if (flags & F_BATTERY) {
        const BatteryCell bat = S.bats[(flags >> F_BAT_SHIFT) - 1];
        float value = bat.emf;
        if (bat.stack >= 0 && S.has_gain[bat.stack]) value *= S.gains[bat.stack];
        if (bat.frequency > 0.0f) {
            const float phase = S.phase_tab[(size_t)(t % TICK_TABLE) * S.n_ac + bat.ac_slot] + bat.frequency * ((float)k * h);
            value *= sinf(6.283185307179586f * phase);
        }
        const int o = (flags >> F_ORI_SHIFT) & 3;
        ex += o == 0 ? value : (o == 2 ? -value : 0.0f);
        ey += o == 1 ? value : (o == 3 ? -value : 0.0f);
    }
    ax = S.qm * (ex + vy * bz);
    ay = S.qm * (ey - vx * bz);
}

// Close-range force of the ions of the conductive cells around an electron at
// fractions (px, py) of its home cell (close_range.rs:335-350). Every cell of
// the window sees the electron at the same fraction of a table sample (cell
// offsets are whole multiples of square_samples), so the bilinear weights are
// computed once; a cell at offset (ox, oy) = home - cell then reads its four
// force samples (the window never reaches the table's edge, where the
// reference clamps).
struct IonWeights {
    int a0, b0, step;
    float w00, w10, w01, w11;
};

__device__ __forceinline__ IonWeights ion_weights(const DevScene &S, float px, float py) {
    const float u = ((px - 0.5f) + S.half) * S.square_samples, v = ((py - 0.5f) + S.half) * S.square_samples;
    const int a = (int)u, b = (int)v;
    const float fu = u - (float)a, fv = v - (float)b;
    return IonWeights{a, b, (int)S.square_samples, (1.0f - fu) * (1.0f - fv), fu * (1.0f - fv), (1.0f - fu) * fv, fu * fv};
}

__device__ __forceinline__ void ion_cell(const DevScene &S, const IonWeights &w, int ox, int oy, float dens, float &fx,
                                         float &fy) {
    const int n = S.square_n;
    const lerp2 *t = S.square_f2 + (size_t)(w.a0 + ox * w.step) * n + (w.b0 + oy * w.step);
    const lerp2 t00 = t[0], t10 = t[n], t01 = t[1], t11 = t[n + 1];
    const float coef = S.k_ion_q * dens;
    fx += coef * (t00.x * w.w00 + t10.x * w.w10 + t01.x * w.w01 + t11.x * w.w11);
    fy += coef * (t00.y * w.w00 + t10.y * w.w10 + t01.y * w.w01 + t11.y * w.w11);
}

// The same summed over a whole window of equal-density cells: every neighbour
// cell shares the electron's fractional table offset, so the sum of their
// bilinear lookups is one lookup in the table summed over the cell offsets.
__device__ __forceinline__ void ion_uniform(const DevScene &S, int home, float px, float py, float &fx, float &fy) {
    const float u = (px - 0.5f + S.half) * S.square_samples, w = (py - 0.5f + S.half) * S.square_samples;
    // A fraction within one FP32 step of 1 rounds u up to the table's last sample
    // (row ION_SUM - 1): start the lerp one sample lower with weight 1 on that
    // row, so the reads stay inside the ION_SUM x ION_SUM table.
    const int a = min(max((int)u - S.ion_a0, 0), ION_SUM - 2), bb = min(max((int)w - S.ion_a0, 0), ION_SUM - 2);
    const float fu = u - (float)(a + S.ion_a0), fv = w - (float)(bb + S.ion_a0);
    const float *tab = S.ion_sum + ((size_t)a * ION_SUM + bb) * 2;
    const float coef = S.k_ion_q * S.density[home];
    const float w00 = (1.0f - fu) * (1.0f - fv), w10 = fu * (1.0f - fv), w01 = (1.0f - fu) * fv, w11 = fu * fv;
    fx += coef * (tab[0] * w00 + tab[ION_SUM * 2] * w10 + tab[2] * w01 + tab[ION_SUM * 2 + 2] * w11);
    fy += coef * (tab[1] * w00 + tab[ION_SUM * 2 + 1] * w10 + tab[3] * w01 + tab[ION_SUM * 2 + 3] * w11);
}

// Cell-mode positions: a neighbour is held as (x fraction within its cell, y
// relative to the window's centre row yc), so its x is the snapshot's FP32
// fraction exactly and |y| < 8 cells resolves 5e-7 cells or better. A run's
// neighbours share a column, so the home electron's x offset from that column
// (fraction + column difference, |x| < 4: 2.4e-7 cells) is formed once per run
// and the pair offset is one FP32 subtraction per axis. An electron and its own
// staged copy give the same bits (its self term is g(0) * 0).
__device__ __forceinline__ float2 cell_pos(int rr, int yc, float fx, float fy) {
    return make_float2(fx, (float)(rr - yc) + fy);
}

// pair_term (below) over FP32 offsets in cells; gs = g_scale turns the squared
// offset into the table coordinate and the sums are in cells (k_pair scales
// them to a force).
__device__ __forceinline__ void pair_term_c(const lerp2 *gt, float gs, float dx, float dy, float &sx, float &sy) {
    const float d2 = (dx * dx + dy * dy) * gs;
    const float tt = fminf(d2, (float)(PAIR_SAMPLES + 1));
    const lerp2 g = gt[(int)tt];
    const float l = g.x + g.y * tt;
    sx += l * dx;
    sy += l * dy;
}

// One pair term g(r^2) d, g lerped from the r^2-uniform table
// (close_range.rs:359-372), each segment held as intercept and slope in the
// table coordinate. The offset (ix, iy) is in 2^-24 cells; s2 =
// g_scale 2^-48 turns its square into the table coordinate r^2 PAIR_SAMPLES /
// cutoff^2. A pair at or beyond the cutoff (d2 >= PAIR_SAMPLES) is clamped to
// entries PAIR_SAMPLES or PAIR_SAMPLES + 1, which are both (0, 0), and the
// electron itself (d = 0) adds g(0) * 0: both contribute exactly zero. The sums
// are in 2^-24 cells (k_pair_fx scales them back).
__device__ __forceinline__ void pair_term(const lerp2 *gt, float s2, int ix, int iy, float &sx, float &sy) {
    const float dx = (float)ix, dy = (float)iy;
    const float d2 = (dx * dx + dy * dy) * s2;
    const float tt = fminf(d2, (float)(PAIR_SAMPLES + 1));
    const lerp2 g = gt[(int)tt];
    const float l = g.x + g.y * tt;
    sx += l * dx;
    sy += l * dy;
}

// kick (physics.rs:57-60), the max_step clamp (physics.rs:151-155) and
// confinement (physics.rs:66-79: x first, then y from the resulting column);
// writes the electron's new cell, fractions and velocity.
__device__ __forceinline__ void finish(const DevScene &S, int id, int hc, int hr, float px, float py, float vx, float vy,
                                       float damping, float ax, float ay, float h) {
    const float scale = 1.0f / (1.0f + damping * h);
    float nvx = (vx + ax * h) * scale, nvy = (vy + ay * h) * scale;
    float mx = nvx * h, my = nvy * h;
    const float len = sqrtf(mx * mx + my * my);
    const float limit = fminf(S.max_step / fmaxf(len, 1e-30f), 1.0f);
    mx *= limit;
    my *= limit;
    float tx = px + mx, ty = py + my;
    int cx = hc, cy = hr;
    resolve(S, 0, hc, hr, tx, nvx);
    renormalise(cx, tx);
    resolve(S, 1, hr, cx, ty, nvy);
    renormalise(cy, ty);
    S.cxy[id] = pack_cell(cx, cy);
    S.fx[id] = tx;
    S.fy[id] = ty;
    S.vx[id] = nvx;
    S.vy[id] = nvy;
}

// CloseRange::force (close_range.rs:314-376) on one electron at fractions
// (px, py) of its home cell (hc, hr), every pair seen from both sides, one
// thread per electron over the sorted snapshot. A window column's cells are
// contiguous in the sorted order, so the column's electrons are walked as one
// run (two cell starts per column, not two per cell; a neighbour's row comes
// with it): a pair at or past the cutoff reads a zero table entry and adds
// exactly nothing, so the sum is the one over the cells within reach. The ion
// terms of the column's conductive cells are summed before its pairs, in
// separate accumulators, in cell order.
__device__ __forceinline__ float2 sparse_close_force(const DevScene &S, const lerp2 *gt, int hc, int hr, float px,
                                                     float py, bool uniform, int home) {
    float fx = 0.0f, fy = 0.0f;
    const int R = S.reach;
    const int c_lo = max(hc - R, 0), c_hi = min(hc + R, S.cols - 1);
    const int r_lo = max(hr - R, 0), r_hi = min(hr + R, S.rows - 1);
    if (uniform) ion_uniform(S, home, px, py, fx, fy);
    const IonWeights iw = ion_weights(S, px, py);
    float sx = 0.0f, sy = 0.0f;
    const int ox = fix24(px), oy = fix24(py);
    for (int cc = c_lo; cc <= c_hi; cc++) {
        const int base = cc * S.rows;
        if (!uniform) {
            for (int rr = r_lo; rr <= r_hi; rr++) {
                const float dens = S.density[base + rr];
                if (dens > 0.0f) ion_cell(S, iw, hc - cc, hr - rr, dens, fx, fy);
            }
        }
        const int jb = S.cstart[base + r_lo], je = S.cstart[base + r_hi + 1];
        const int dx = (hc - cc) * (1 << 24) + ox;
        for (int j = jb; j < je; j++) {
            const int rr = S.scell[j] - base;
            const float2 f = snap_pos(S, j);
            pair_term(gt, S.pair_s2, dx - fix24(f.x), (hr - rr) * (1 << 24) + oy - fix24(f.y), sx, sy);
        }
    }
    return make_float2(fx + S.k_pair_fx * sx, fy + S.k_pair_fx * sy);
}

// One electron's substep (physics.rs:139-158), one thread per electron: `slot`
// is its place in the sorted snapshot. Inlined into k_push_sparse: as a call,
// each electron saved and restored the caller's registers through scratch
// (420 MB written and read back per launch at 4096^2, 5M electrons).
__device__ __forceinline__ void push_core(const DevScene &S, const lerp2 *gt, int slot, int k, float h, uint32_t t) {
    const int id = S.order[slot];
    const int home = S.scell[slot];
    const int hc = home / S.rows, hr = home - (home / S.rows) * S.rows;
    const float4 e4 = S.snap[slot];
    const float px = e4.x, py = e4.y;   // fractions within the home cell
    const float vx = e4.z, vy = e4.w;
    const CellInfo info = S.info[home];
    float ax, ay;
    field_accel(S, hc, hr, px, py, vx, vy, info.flags, k, h, t, ax, ay);
    if (S.close_on) {
        const float2 f = sparse_close_force(S, gt, hc, hr, px, py, (info.flags & F_ION_UNIFORM) != 0u, home);
        ax += f.x * S.inv_mass;
        ay += f.y * S.inv_mass;
    }
    finish(S, id, hc, hr, px, py, vx, vy, info.damping, ax, ay, h);
}

// The substep of one electron whose close-range + ion force f is already known.
__device__ __forceinline__ void kick_core(const DevScene &S, int slot, int k, float h, uint32_t t, float2 f) {
    const int id = S.order[slot];
    const int home = S.scell[slot];
    const int hc = home / S.rows, hr = home - (home / S.rows) * S.rows;
    const float4 e4 = S.snap[slot];
    const float px = e4.x, py = e4.y;
    const float vx = e4.z, vy = e4.w;
    const CellInfo info = S.info[home];
    float ax, ay;
    field_accel(S, hc, hr, px, py, vx, vy, info.flags, k, h, t, ax, ay);
    ax += f.x * S.inv_mass;
    ay += f.y * S.inv_mass;
    finish(S, id, hc, hr, px, py, vx, vy, info.damping, ax, ay, h);
}

// (a call from k_rest's queue, whose tile items hold many registers already)
__device__ __noinline__ void push_one(const DevScene &S, int slot, int k, float h, uint32_t t) {
    push_core(S, S.g_lerp, slot, k, h, t);
}

// Lanes per electron in the cell mode: the cell's count rounded up to a power of
// two, at most a wave.
__device__ __forceinline__ int lanes_per(int n) {
    int i = 1;
    while (i < n && i < 64) i <<= 1;
    return i;
}

// Column runs of the cell mode: the electrons of rows r_lo..r_hi of column cc,
// contiguous in the sorted order, with cell_pos positions (yc = wr0 + W / 2).
// Staged in LDS, or read from the snapshot.
struct LdsRuns {
    const float2 *lp;
    const int *lb, *le;
    int wc0, wr0, W;
    __device__ void run(int cc, int r_lo, int r_hi, int &jb, int &je) const {
        const int w = (cc - wc0) * W;
        jb = lb[w + r_lo - wr0];
        je = le[w + r_hi - wr0];
    }
    __device__ float2 at(int j, int) const { return lp[j]; }
};

struct GlobalRuns {
    const DevScene &S;
    int wc0, wr0;
    __device__ void run(int cc, int r_lo, int r_hi, int &jb, int &je) const {
        jb = S.cstart[cc * S.rows + r_lo];
        je = S.cstart[cc * S.rows + r_hi + 1];
    }
    int yc;
    __device__ float2 at(int j, int cc) const {
        const int rr = S.scell[j] - cc * S.rows;
        const float2 f = snap_pos(S, j);
        return cell_pos(rr, yc, f.x, f.y);
    }
};

// Lanes per electron of a cell-mode work item over a cell of n electrons, E
// electrons per lane: ceil(n / E) rounded up to a power of two, at most imax.
__device__ __forceinline__ int item_lanes(int n, int E, int imax) { return min(lanes_per((n + E - 1) / E), imax); }

// What a cell-mode work item does with an electron's close-range + ion force:
// push the electron (ITEM_PUSH), write the force to force_out[id] (ITEM_FORCE),
// or store it by sorted slot in S.pforce for k_push_sparse to push (ITEM_STORE).
enum { ITEM_PUSH = 0, ITEM_FORCE = 1, ITEM_STORE = 2 };

// One work item of the cell mode: the E I electrons [chunk E I, chunk E I + E I)
// of home cell (hc, hr), I = item_lanes(n, E, imax); lane i + I sl holds
// electrons chunk E I + e I + i (e < E) and walks slice sl of SL = 64 / I slices
// of the neighbour runs, so each neighbour read from LDS serves E pair terms.
// The partial sums meet in a fixed butterfly; lane sl = e then takes electron e
// as MODE says.
template <int MODE, int E, class RUNS>
__device__ __forceinline__ void cell_item(const DevScene &S, const RUNS &runs, const lerp2 *gt, const float *dens_w,
                                          int hc, int hr, int e0, int n, int chunk, int wc0, int wr0, int W, int k,
                                          float h, uint32_t t, float2 *force_out, int imax = 64) {
    static_assert(E == 1 || E == 2, "one or two electrons per lane");
    const int lane = threadIdx.x & 63;
    const int R = S.reach;
    const int home = hc * S.rows + hr;
    // Every chunk but a cell's last holds E Ifull electrons; the last takes as
    // many lanes as its own electrons need.
    // (read as wave-uniform: the loop bounds below derive from I)
    const int imx = E == 2 ? min(imax, 32) : imax;
    const int base = __builtin_amdgcn_readfirstlane(chunk * E * item_lanes(n, E, imx));
    const int I = __builtin_amdgcn_readfirstlane(item_lanes(n - base, E, imx)), shift = __builtin_ctz(I), SL = 64 >> shift;
    const int i = lane & (I - 1), sl = lane >> shift;
    int slot[E];
    bool live[E];
    const float gs = S.g_scale;
    const int yc = wr0 + (W >> 1);
    float px[E], by[E];
    float sx[E], sy[E];
#pragma unroll
    for (int e = 0; e < E; e++) {
        const int ei = base + e * I + i;
        live[e] = ei < n;
        slot[e] = e0 + (live[e] ? ei : n - 1);
        const float2 f = snap_pos(S, slot[e]);
        px[e] = f.x;
        by[e] = cell_pos(hr, yc, f.x, f.y).y;
        sx[e] = 0.0f;
        sy[e] = 0.0f;
    }
    const int c_lo = max(hc - R, 0), c_hi = min(hc + R, S.cols - 1);
    const int r_lo = max(hr - R, 0), r_hi = min(hr + R, S.rows - 1);
    for (int cc = c_lo; cc <= c_hi; cc++) {
        int jb, je;
        runs.run(cc, r_lo, r_hi, jb, je);
        jb = __builtin_amdgcn_readfirstlane(jb);
        je = __builtin_amdgcn_readfirstlane(je);
        float bx[E];   // x offset of the home electrons from column cc's cell corners
#pragma unroll
        for (int e = 0; e < E; e++) bx[e] = px[e] + (float)(hc - cc);
        int jj = jb;
        for (; jj + 2 * SL <= je; jj += 2 * SL) {
            const float2 q0 = runs.at(jj + sl, cc), q1 = runs.at(jj + SL + sl, cc);
#pragma unroll
            for (int e = 0; e < E; e++) {
                pair_term_c(gt, gs, bx[e] - q0.x, by[e] - q0.y, sx[e], sy[e]);
                pair_term_c(gt, gs, bx[e] - q1.x, by[e] - q1.y, sx[e], sy[e]);
            }
        }
        for (; jj < je; jj += SL) {
            // past the run's end a lane contributes nothing
            const float2 q = runs.at(min(jj + sl, je - 1), cc);
            const bool in = jj + sl < je;
#pragma unroll
            for (int e = 0; e < E; e++)
                pair_term_c(gt, gs, in ? bx[e] - q.x : 0.0f, in ? by[e] - q.y : 0.0f, sx[e], sy[e]);
        }
    }
    // From here on a lane works for electron e = sl % E (the ion terms) and, if
    // sl < E, finishes it.
    const int me = E == 2 ? (sl & 1) : 0;
    // (the y fraction is read again rather than kept through the pair loop)
    const float mpx = E == 2 && me ? px[E - 1] : px[0], mpy = snap_pos(S, E == 2 && me ? slot[E - 1] : slot[0]).y;
    const CellInfo info = S.info[home];
    const bool uniform = (info.flags & F_ION_UNIFORM) != 0u;
    float fx = 0.0f, fy = 0.0f;
    if (!uniform) {
        const IonWeights iw = ion_weights(S, mpx, mpy);
        const int nr = r_hi - r_lo + 1, cells = (c_hi - c_lo + 1) * nr;
        if (cells <= 64) {
            // The window's cells that hold ions, compacted in cell order: lane w
            // tests cell w, and the k-th cell with ions goes to lane k (a forward
            // permute; the empty cells take the lanes after them). Lanes then
            // walk the list, not the whole window: near a conductor's edge most
            // window cells are empty.
            const int cc = c_lo + lane / nr, rr = r_lo + lane % nr;
            const float dens = lane < cells ? dens_w[(cc - wc0) * W + (rr - wr0)] : 0.0f;
            const unsigned long long m = __ballot(dens > 0.0f);
            const int K = __popcll(m);
            const int below = __builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
            const int dst = dens > 0.0f ? below : K + (lane - below);
            const int packed = ((hc - cc + 8) << 4) | (hr - rr + 8);   // offsets in [-4, 4]
            const int cpk = __builtin_amdgcn_ds_permute(dst * 4, packed);
            const int cdn = __builtin_amdgcn_ds_permute(dst * 4, __float_as_int(dens));
            // (the loop bound is the wave's: a lane reads its entry from another
            // lane, which must still be active, so every lane runs every round)
            for (int j0 = 0; j0 < K; j0 += SL / E) {
                const int j = j0 + sl / E;
                const int pk = __builtin_amdgcn_ds_bpermute(min(j, 63) * 4, cpk);
                const float dn = __int_as_float(__builtin_amdgcn_ds_bpermute(min(j, 63) * 4, cdn));
                if (j < K) ion_cell(S, iw, (pk >> 4) - 8, (pk & 15) - 8, dn, fx, fy);
            }
        } else {
            for (int w = sl / E; w < cells; w += SL / E) {
                const int cc = c_lo + w / nr, rr = r_lo + w % nr;
                const float dens = dens_w[(cc - wc0) * W + (rr - wr0)];
                if (dens > 0.0f) ion_cell(S, iw, hc - cc, hr - rr, dens, fx, fy);
            }
        }
    }
    for (int off = I; off < 64; off <<= 1) {
#pragma unroll
        for (int e = 0; e < E; e++) {
            sx[e] += __shfl_xor(sx[e], off, 64);
            sy[e] += __shfl_xor(sy[e], off, 64);
        }
        if (off >= E * I) {   // the ion sums stay apart per electron
            fx += __shfl_xor(fx, off, 64);
            fy += __shfl_xor(fy, off, 64);
        }
    }
    const bool mlive = E == 2 && me ? live[E - 1] : live[0];
    if (sl < E && mlive) {
        const int mslot = E == 2 && me ? slot[E - 1] : slot[0];
        if (uniform) ion_uniform(S, home, mpx, mpy, fx, fy);
        fx += S.k_pair * (E == 2 && me ? sx[E - 1] : sx[0]);
        fy += S.k_pair * (E == 2 && me ? sy[E - 1] : sy[0]);
        if (MODE == ITEM_FORCE) {
            force_out[S.order[mslot]] = make_float2(fx, fy);
        } else if (MODE == ITEM_STORE) {
            S.pforce[mslot] = make_float2(fx, fy);
        } else {
            const float2 v = snap_vel(S, mslot);
            const float vx = v.x, vy = v.y;
            float ax, ay;
            field_accel(S, hc, hr, mpx, mpy, vx, vy, info.flags, k, h, t, ax, ay);
            ax += fx * S.inv_mass;
            ay += fy * S.inv_mass;
            finish(S, S.order[mslot], hc, hr, mpx, mpy, vx, vy, info.damping, ax, ay, h);
        }
    }
}

// Lanes per electron pair of a cell-mode item at most: a cell of more than
// 2 CELL_IMAX electrons splits into items of 2 CELL_IMAX and a last one sized to
// the rest (17 electrons: 8 + 1 lanes' walks, not 16).
constexpr int CELL_IMAX = 8;

// Electrons per lane in the cell mode: two share each neighbour read, except in
// a cell of one.
__device__ __forceinline__ int item_e(int n) { return n >= 2 ? 2 : 1; }

// The close-range force alone on one electron (electron mode of FORCE).
__device__ void force_core(const DevScene &S, const lerp2 *gt, int slot, float2 *force_out) {
    const int home = S.scell[slot];
    const int hc = home / S.rows, hr = home - (home / S.rows) * S.rows;
    const float2 f = snap_pos(S, slot);
    const bool uniform = (S.info[home].flags & F_ION_UNIFORM) != 0u;
    force_out[S.order[slot]] = sparse_close_force(S, gt, hc, hr, f.x, f.y, uniform, home);
}

// Substep k for the electrons of the crowded push tiles (tile_dense), per
// PUSH_TILE x PUSH_TILE tile of cells (FORCE: the close-range force alone,
// CloseRange::force, close_range.rs:314-376). The positions of the tile's
// neighbourhood (reach cells around it) are staged in LDS, in fixed point
// relative to the window's corner, when they fit (else read from the snapshot).
// The pair sums run in cell mode: a wave takes up to I electrons of one home
// cell, I a power of two, and splits the window's (2 reach + 1) column runs
// between 64 / I lanes per electron, so every lane of the wave walks the same
// neighbour list without divergence; the partial sums meet in a fixed
// butterfly. The other tiles' electrons go through k_push_sparse. Every sum
// runs in a fixed order, so a scene gives the same bits alone or batched.
//
// The tick's k_push_tiled runs 16 waves per block: two blocks fill all 32 wave
// slots of a CU (8 per SIMD), which the pair loop's LDS and dependency latency
// needs (the kernel is held to the 64 VGPRs that allows, with a few spills
// outside the pair loop): 0.61 ms against 0.70 ms with 8 waves on the 512^2,
// 16-carrier scene. The force pass k_force_tiled keeps 8 waves per block and
// up to 128 VGPRs, which measured faster on the clustered 4096^2 scene (2.31
// against 2.56 ms), where cells of up to 63 electrons make long items.
constexpr int PUSH_WIN = PUSH_TILE + 2 * 4;   // window side for reach <= 4
constexpr int PUSH_STAGE = 5120;   // positions (40 KB): with the 33 KB pair table, under 80 KB: two blocks per CU
constexpr int PUSH_CELLS = PUSH_TILE * PUSH_TILE;
// The dense-tile list's last tiles, up to PUSH_SPLIT_TILES of them and as many
// as the scene has push blocks beyond its dense tiles, are pushed in PUSH_SPLIT
// parts (a divisor of PUSH_TILE), each part by a block of its own: the blocks
// that start last, and set the launch's end, hold half a tile. On the 512^2
// scene: 0.498 -> 0.478 ms per tick with 256 (128: 0.485, 512: 0.486, every
// spare block: 0.489; four parts of 128 or 256 tiles: 0.484, 0.483), the same
// bits; the README grids and the clustered force pass unchanged.
constexpr int PUSH_SPLIT = 2;
constexpr int PUSH_SPLIT_TILES = 256;
constexpr int PUSH_THREADS = 1024;
constexpr int FORCE_THREADS = 512;

// LDS of the cell-mode tile body (the positions, the pair table, the window's
// runs and densities, the tile's work items).
struct PushTileLds {
    float2 lp[PUSH_STAGE];
    lerp2 gt[PAIR_SAMPLES + 4];
    int lb[PUSH_WIN * PUSH_WIN], le[PUSH_WIN * PUSH_WIN];
    float dens_w[PUSH_WIN * PUSH_WIN];               // ion density of the window's cells
    int cell_e0[PUSH_CELLS], cell_n[PUSH_CELLS];     // tile cells: first sorted slot, electrons
    int col_base[PUSH_WIN + 1], col_g0[PUSH_WIN];   // LDS base and first global slot per column
    int item_pref[PUSH_CELLS + 1];                   // cell mode: prefix of work items per tile cell
    int item_next;                                   // the pass's next work item, drawn by the waves
};

// One push tile in cell mode (every thread of the block; the pair table must be
// in L.gt). With HEAVY, cells of HEAVY_CELL or more electrons are pushed here
// too (the tick leaves them to k_push_heavy).
template <int MODE, int THREADS, bool HEAVY>
__device__ __forceinline__ void push_tile(const DevScene &S, PushTileLds &L, int local, int k, float h, uint32_t t,
                                          float2 *force_out, int a_lo = 0, int a_hi = PUSH_TILE) {
    constexpr int WAVES = THREADS / 64;
    float2 *lp = L.lp;
    const lerp2 *gt = L.gt;
    int *lb = L.lb, *le = L.le, *cell_e0 = L.cell_e0, *cell_n = L.cell_n, *col_base = L.col_base, *col_g0 = L.col_g0;
    int *item_pref = L.item_pref;
    float *dens_w = L.dens_w;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int tiles_r = (S.rows + PUSH_TILE - 1) / PUSH_TILE;
    const int tc0 = (local / tiles_r) * PUSH_TILE, tr0 = (local % tiles_r) * PUSH_TILE;
    const int tc1 = min(tc0 + PUSH_TILE, S.cols), tr1 = min(tr0 + PUSH_TILE, S.rows);
    const int R = S.close_on ? S.reach : 0;
    const int W = PUSH_TILE + 2 * R;
    const int wc0 = tc0 - R, wr0 = tr0 - R;
    const int rlo = max(wr0, 0), rhi = min(wr0 + W, S.rows) - 1;
    // Wave 0 lays out the window (column runs, lane = column) and the tile's
    // work (lane = tile cell, column-major).
    if (wave == 0) {
        const int cc = wc0 + lane;
        int g0 = 0, n = 0;
        if (lane < W && cc >= 0 && cc < S.cols && rlo <= rhi) {
            g0 = S.cstart[cc * S.rows + rlo];
            n = S.cstart[cc * S.rows + rhi + 1] - g0;
        }
        const int incl = wave_scan_incl(n);
        if (lane < W) {
            col_base[lane] = incl - n;
            col_g0[lane] = g0;
        }
        if (lane == W - 1) col_base[W] = incl;
        const int hc = tc0 + lane / PUSH_TILE, hr = tr0 + lane % PUSH_TILE;
        int m = 0, first = 0;
        if (hc < tc1 && hr < tr1) {
            const int cell = hc * S.rows + hr;
            first = S.cstart[cell];
            m = S.cstart[cell + 1] - first;
        }
        cell_e0[lane] = first;
        cell_n[lane] = m;
        const int E = item_e(m), I = item_lanes(m, E, E == 2 ? CELL_IMAX : 64);
        const int items = (m > 0 && (HEAVY || m < HEAVY_CELL)) ? (m + E * I - 1) / (E * I) : 0;
        const int item_incl = wave_scan_incl(items);
        item_pref[lane] = item_incl - items;
        if (lane == 63) item_pref[PUSH_CELLS] = item_incl;
    }
    __syncthreads();
    for (int q = threadIdx.x; q < W * W; q += THREADS) {
        const int ci = q / W, ri = q - (q / W) * W;
        const int cc = wc0 + ci, rr = wr0 + ri;
        const bool in = cc >= 0 && cc < S.cols && rr >= 0 && rr < S.rows;
        dens_w[q] = in ? S.density[cc * S.rows + rr] : 0.0f;
    }
    // As many consecutive columns of home cells per pass as the staging holds
    // (home columns [a, b) need window columns [a, b + 2 reach)): the whole
    // window in one pass when it fits; a single column that does not fit
    // reads its neighbours from global memory.
    for (int a = a_lo, b = a_lo; a < a_hi; a = b) {
        b = a + 1;
        while (b < a_hi && col_base[min(b + 1 + 2 * R, W)] - col_base[a] <= PUSH_STAGE) b++;
        const int ca = a, cb = min(b + 2 * R, W);
        const int la = a * PUSH_TILE, lz = b * PUSH_TILE;
        const int base = col_base[ca], staged_n = col_base[cb] - base;
        const bool staged = staged_n <= PUSH_STAGE;
        if (threadIdx.x == 0) L.item_next = item_pref[la];
        if (staged) {
            for (int q = threadIdx.x; q < (cb - ca) * W; q += THREADS) {
                const int ci = ca + q / W, ri = q - (q / W) * W;
                const int cc = wc0 + ci, rr = wr0 + ri;
                const int w = ci * W + ri;
                if (cc < 0 || cc >= S.cols || rr < 0 || rr >= S.rows) {
                    lb[w] = 0;
                    le[w] = 0;
                } else {
                    const int cell = cc * S.rows + rr;
                    lb[w] = col_base[ci] - base + (S.cstart[cell] - col_g0[ci]);
                    le[w] = col_base[ci] - base + (S.cstart[cell + 1] - col_g0[ci]);
                }
            }
            // One pass over every staged electron (column found by search), so a
            // sparse window costs one round of loads rather than one per column.
            for (int j = threadIdx.x; j < staged_n; j += THREADS) {
                int q = ca;
                for (int step = PUSH_WIN / 2; step > 0; step >>= 1)
                    if (q + step < cb && col_base[q + step] - base <= j) q += step;
                const int g = col_g0[q] + (j - (col_base[q] - base));
                const int rr = S.scell[g] - (wc0 + q) * S.rows;
                const float2 f = snap_pos(S, g);
                lp[j] = cell_pos(rr, wr0 + (W >> 1), f.x, f.y);
            }
        }
        __syncthreads();
        // Each wave draws its next item from an LDS counter as it finishes the
        // last, so the waves end the pass together although items differ in
        // cost (a cell's last item is short); any wave computes an item alike.
        const int it_end = item_pref[lz];
        for (;;) {
            int it = 0;
            if (lane == 0) it = atomicAdd(&L.item_next, 1);
            it = __builtin_amdgcn_readfirstlane(it);
            if (it >= it_end) break;
            int lo = la, hi = lz;
            while (hi - lo > 1) {
                const int mid = (lo + hi) >> 1;
                if (item_pref[mid] <= it) lo = mid;
                else hi = mid;
            }
            lo = __builtin_amdgcn_readfirstlane(lo);
            const int hc = tc0 + lo / PUSH_TILE, hr = tr0 + lo % PUSH_TILE;
            const int chunk = it - item_pref[lo];
            const int e0 = __builtin_amdgcn_readfirstlane(cell_e0[lo]), n = __builtin_amdgcn_readfirstlane(cell_n[lo]);
            if (staged) {
                const LdsRuns runs{lp, lb, le, wc0, wr0, W};
                if (item_e(n) == 2)
                    cell_item<MODE, 2>(S, runs, gt, dens_w, hc, hr, e0, n, chunk, wc0, wr0, W, k, h, t, force_out, CELL_IMAX);
                else
                    cell_item<MODE, 1>(S, runs, gt, dens_w, hc, hr, e0, n, chunk, wc0, wr0, W, k, h, t, force_out, CELL_IMAX);
            } else {
                const GlobalRuns runs{S, wc0, wr0, wr0 + (W >> 1)};
                if (item_e(n) == 2)
                    cell_item<MODE, 2>(S, runs, gt, dens_w, hc, hr, e0, n, chunk, wc0, wr0, W, k, h, t, force_out, CELL_IMAX);
                else
                    cell_item<MODE, 1>(S, runs, gt, dens_w, hc, hr, e0, n, chunk, wc0, wr0, W, k, h, t, force_out, CELL_IMAX);
            }
        }
        __syncthreads();
    }
}

template <int MODE, int THREADS>
__device__ __forceinline__ void push_tiled_body(Batch b, int k, float2 *force_out, int first, int stride) {
    constexpr bool FORCE = MODE == ITEM_FORCE;
    __shared__ PushTileLds L;
    bool table_loaded = false;
    for (int blk = first; blk < b.total[KC_PUSH]; blk += stride) {
        const int s = scene_of(b, KC_PUSH, blk);
        const DevScene &S = b.scenes[s];
        const uint32_t t = S.ctr->tick;
        int sub = 1;
        if (!FORCE) {
            sub = substeps_of(S, t);
            if (k >= sub) continue;
        } else if (!S.close_on) {
            continue;
        }
        const float h = S.dt / (float)sub;
        const int local = block_in_scene(b, KC_PUSH, s, blk);
        if (!FORCE && k == 0 && threadIdx.x == 0 && local == 0) S.ctr->substeps = sub;
        if (!S.close_on) continue;
        // The scene's blocks share out work units in turn: the sort's list of
        // dense tiles, whole, then its last PUSH_SPLIT_TILES tiles in parts of
        // PUSH_TILE / PUSH_SPLIT home columns. A part stages the window columns
        // its home columns reach and computes every electron as the whole tile
        // would, so the results do not depend on the split.
        const int *pre = b.prefix + KC_PUSH * (b.n_scenes + 1);
        const int nblk = pre[s + 1] - pre[s], dense_n = S.ctr->dense_n;
        // (as many tiles as the scene has blocks beyond its dense tiles, so that
        // each unit has a block of its own)
        const int split = min(min(dense_n, PUSH_SPLIT_TILES), max(nblk - dense_n, 0)), full = dense_n - split;
        const int units = full + split * PUSH_SPLIT;
        for (int u = local; u < units; u += nblk) {
            if (!table_loaded) {   // the pair table is shared by every scene
                for (int j = threadIdx.x; j < PAIR_SAMPLES + 4; j += THREADS) L.gt[j] = S.g_lerp[j];
                table_loaded = true;
            }
            const bool whole = u < full;
            const int q = whole ? u : full + (u - full) / PUSH_SPLIT, part = whole ? 0 : (u - full) % PUSH_SPLIT;
            const int a_lo = whole ? 0 : part * (PUSH_TILE / PUSH_SPLIT);
            const int a_hi = whole ? PUSH_TILE : a_lo + PUSH_TILE / PUSH_SPLIT;
            push_tile<MODE, THREADS, false>(S, L, S.dense_tiles[q], k, h, t, force_out, a_lo, a_hi);
        }
    }
}

// Substep k (FORCE: the close-range force) for the electrons of the other push
// tiles, one thread per electron in sorted order, neighbours read from the
// snapshot in global memory with per-cell culling against the cutoff. With few
// electrons per cell there is little to share between neighbours, and a tile
// would be a chain of dependent loads for a handful of threads.
template <bool FORCE>
// (held to 64 VGPRs, 8 waves per SIMD: its gathers are latency-bound; 7 us
// faster on the 512^2 scene than at the 100 VGPRs it would take)
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8))) void k_push_sparse(Batch b, int k,
                                                                                             float2 *force_out) {
    for (int blk = blockIdx.x; blk < b.total[KC_ELECTRONS]; blk += gridDim.x) {
        const int s = scene_of(b, KC_ELECTRONS, blk);
        const DevScene &S = b.scenes[s];
        const uint32_t t = S.ctr->tick;
        int sub = 1;
        if (!FORCE) {
            sub = substeps_of(S, t);
            if (k >= sub) continue;
        } else if (!S.close_on) {
            continue;
        }
        const float h = S.dt / (float)sub;
        const int slot = block_in_scene(b, KC_ELECTRONS, s, blk) * 256 + (int)threadIdx.x;
        if (!FORCE && k == 0 && slot == 0) S.ctr->substeps = sub;
        if (slot >= S.n) continue;
        const int cell = S.scell[slot];
        const int cc = cell / S.rows, rr = cell - (cell / S.rows) * S.rows;
        if (S.close_on && S.tile_dense[(cc / PUSH_TILE) * ((S.rows + PUSH_TILE - 1) / PUSH_TILE) + rr / PUSH_TILE]) {
            // k_push_tiled stored the close-range force of this electron unless its
            // cell is k_push_heavy's; the field kick and move run here, one lane
            // per electron.
            if (!FORCE && S.cstart[cell + 1] - S.cstart[cell] < HEAVY_CELL) kick_core(S, slot, k, h, t, S.pforce[slot]);
            continue;
        }
        if (FORCE)
            force_core(S, S.g_lerp, slot, force_out);
        else
            push_core(S, S.g_lerp, slot, k, h, t);
    }
}

template __global__ void k_push_sparse<false>(Batch, int, float2 *);
template __global__ void k_push_sparse<true>(Batch, int, float2 *);

// Substep k (FORCE: the close-range force) for the electrons of crowded cells
// (>= HEAVY_CELL, listed by the sort). A block takes whole cells from the list,
// stages the cell's (2 reach + 1)^2 window and spreads the cell's 32-electron
// chunks (16 lanes per electron pair) over its waves: a cluster's core is shared
// among many CUs instead of weighing on its tile's one block. Blocks of 8 waves:
// a cell of a few hundred electrons has about that many chunks.
constexpr int HEAVY_I = 16;
constexpr int HEAVY_WIN = 2 * 4 + 1;   // window side for reach <= 4
constexpr int HEAVY_THREADS = 512;
constexpr int HEAVY_WAVES = HEAVY_THREADS / 64;

template <bool FORCE>
__global__ __launch_bounds__(HEAVY_THREADS) void k_push_heavy(Batch b, int k, float2 *force_out) {
    __shared__ float2 lp[PUSH_STAGE];
    __shared__ lerp2 gt[PAIR_SAMPLES + 4];
    __shared__ int lb[HEAVY_WIN * HEAVY_WIN], le[HEAVY_WIN * HEAVY_WIN];
    __shared__ float dens_w[HEAVY_WIN * HEAVY_WIN];
    __shared__ int col_base[HEAVY_WIN + 1], col_g0[HEAVY_WIN];
    bool table_loaded = false;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int blk = blockIdx.x; blk < b.total[KC_HEAVY]; blk += gridDim.x) {
        const int s = scene_of(b, KC_HEAVY, blk);
        const DevScene &S = b.scenes[s];
        if (!S.close_on) continue;
        const uint32_t t = S.ctr->tick;
        int sub = 1;
        if (!FORCE) {
            sub = substeps_of(S, t);
            if (k >= sub) continue;
        }
        const float h = S.dt / (float)sub;
        const int *pre = b.prefix + KC_HEAVY * (b.n_scenes + 1);
        const int nblk = pre[s + 1] - pre[s];
        const int heavy_n = S.ctr->heavy_n;
        for (int hcell = block_in_scene(b, KC_HEAVY, s, blk); hcell < heavy_n; hcell += nblk) {
            if (!table_loaded) {   // the pair table is shared by every scene
                for (int q = threadIdx.x; q < PAIR_SAMPLES + 4; q += HEAVY_THREADS) gt[q] = S.g_lerp[q];
                table_loaded = true;
            }
            const int cell = S.heavy_cells[hcell];
            const int hc = cell / S.rows, hr = cell - (cell / S.rows) * S.rows;
            const int R = S.reach, W = 2 * R + 1;
            const int wc0 = hc - R, wr0 = hr - R;
            const int rlo = max(wr0, 0), rhi = min(wr0 + W, S.rows) - 1;
            if (wave == 0) {
                const int cc = wc0 + lane;
                int g0 = 0, m = 0;
                if (lane < W && cc >= 0 && cc < S.cols) {
                    g0 = S.cstart[cc * S.rows + rlo];
                    m = S.cstart[cc * S.rows + rhi + 1] - g0;
                }
                const int incl = wave_scan_incl(m);
                if (lane < W) {
                    col_base[lane] = incl - m;
                    col_g0[lane] = g0;
                }
                if (lane == W - 1) col_base[W] = incl;
            }
            __syncthreads();
            const bool staged = col_base[W] <= PUSH_STAGE;
            for (int q = threadIdx.x; q < W * W; q += HEAVY_THREADS) {
                const int ci = q / W, ri = q - (q / W) * W;
                const int cc = wc0 + ci, rr = wr0 + ri;
                const bool in = cc >= 0 && cc < S.cols && rr >= 0 && rr < S.rows;
                const int c2 = cc * S.rows + rr;
                dens_w[q] = in ? S.density[c2] : 0.0f;
                if (staged) {
                    lb[q] = in ? col_base[ci] + (S.cstart[c2] - col_g0[ci]) : 0;
                    le[q] = in ? col_base[ci] + (S.cstart[c2 + 1] - col_g0[ci]) : 0;
                }
            }
            if (staged) {
                const int staged_n = col_base[W];
                for (int j = threadIdx.x; j < staged_n; j += HEAVY_THREADS) {
                    int q = 0;
                    for (int step = 8; step > 0; step >>= 1)   // W <= HEAVY_WIN = 9
                        if (q + step < W && col_base[q + step] <= j) q += step;
                    const int g = col_g0[q] + (j - col_base[q]);
                    const int rr = S.scell[g] - (wc0 + q) * S.rows;
                    const float2 f = snap_pos(S, g);
                    lp[j] = cell_pos(rr, wr0 + (W >> 1), f.x, f.y);
                }
            }
            __syncthreads();
            const int e0 = __builtin_amdgcn_readfirstlane(S.cstart[cell]);
            const int n = __builtin_amdgcn_readfirstlane(S.cstart[cell + 1]) - e0;
            const int chunks = (n + 2 * HEAVY_I - 1) / (2 * HEAVY_I);   // n >= HEAVY_CELL: I = HEAVY_I
            for (int ch = wave; ch < chunks; ch += HEAVY_WAVES) {
                constexpr int MODE = FORCE ? ITEM_FORCE : ITEM_PUSH;
                if (staged)
                    cell_item<MODE, 2>(S, LdsRuns{lp, lb, le, wc0, wr0, W}, gt, dens_w, hc, hr, e0, n, ch, wc0, wr0, W, k,
                                       h, t, force_out, HEAVY_I);
                else
                    cell_item<MODE, 2>(S, GlobalRuns{S, wc0, wr0, wr0 + (W >> 1)}, gt, dens_w, hc, hr, e0, n, ch, wc0, wr0, W, k, h, t,
                                       force_out, HEAVY_I);
            }
            __syncthreads();
        }
    }
}

template __global__ void k_push_heavy<false>(Batch, int, float2 *);
template __global__ void k_push_heavy<true>(Batch, int, float2 *);

// The trim alone (a batch with more than TRIM_FUSE_MAX trim blocks): the same
// blocks of PUSH_THREADS threads, so a scene's gains have the same bits either way.
constexpr int TRIM_FUSE_MAX = 256;
__global__ __launch_bounds__(PUSH_THREADS) void k_trim(Batch b) {
    for (int blk = blockIdx.x; blk < b.total[KC_STACKS]; blk += gridDim.x) trim_block<PUSH_THREADS>(b, blk);
}

// The tick's first kernel: blocks [0, trim_grid) compute the battery trim's
// terminal potentials (trim_block), the others run the tiled push of substep k.
// The trim reads only what the tick before left (the potential and the sorted
// snapshot) and the push does not read the gains, so the trim's latency-bound
// blocks overlap the push instead of running before it.
__global__ __launch_bounds__(PUSH_THREADS) __attribute__((amdgpu_waves_per_eu(8))) void k_push_tiled(Batch b, int k,
                                                                                                  int trim_grid) {
    const int push_grid = (int)gridDim.x - trim_grid;
    if ((int)blockIdx.x >= push_grid) {
        for (int blk = (int)blockIdx.x - push_grid; blk < b.total[KC_STACKS]; blk += trim_grid) trim_block<PUSH_THREADS>(b, blk);
        return;
    }
    push_tiled_body<ITEM_STORE, PUSH_THREADS>(b, k, nullptr, (int)blockIdx.x, push_grid);
}

__global__ __launch_bounds__(FORCE_THREADS) void k_force_tiled(Batch b, float2 *force_out) {
    push_tiled_body<ITEM_FORCE, FORCE_THREADS>(b, 0, force_out, (int)blockIdx.x, (int)gridDim.x);
}

// Substeps 1.. of a tick that needs several (fast electrons; physics.rs:124-162:
// every substep takes its forces from the positions at its start, so the cell
// list is rebuilt before each, sim.rs:643-651). The host queues ticks without
// knowing the substep count, so one launch per tick covers whatever the tick
// needs, as a work queue: the blocks of a scene draw numbered items from the
// scene's ticket counter, the items of each substep in phase order
//   0 count      1024 electrons: cell and rank within it (atomic count)
//   1 tile sum   SCAN_TILE cells: their electron count
//   2 top scan   one item: exclusive scan of the tile sums, cstart[cells] = n,
//                dense-tile flags cleared
//   3 starts     SCAN_TILE cells: cstart from the scan, counts back to zero
//   4 scatter    1024 electrons: order[cstart + rank] = electron
//   5 canon      REST_THREADS cells: in-cell order by electron index, the
//                snapshot, the dense-tile flags
//   6 push       one push tile: a dense tile in cell mode (push_tile, crowded
//                cells included), the others electron by electron (push_one)
// and an item first waits until every item of the phase before it is done.
// Those items were drawn earlier, by blocks already running, whose own waits
// go back further in the same way to the first phase, which waits for nothing:
// so every wait ends, whatever number of blocks is resident at once. When no
// scene needs a second substep every block leaves after one read.
//
// A wait that runs past REST_SPIN_LIMIT polls (about a second) marks the scene's
// rest_error and drops the rest of the tick's substeps for that block, so a
// fault in this bookkeeping shows as an error instead of a hung GPU.
constexpr int REST_SPIN_LIMIT = 1 << 24;

__device__ __forceinline__ int rest_items(int p, int n, int cells, int ptiles) {
    switch (p) {
    case 0: case 4: return (n + 1023) / 1024;
    case 1: case 3: return (cells + SCAN_TILE - 1) / SCAN_TILE;
    case 2: return 1;
    case 5: return (cells + REST_THREADS - 1) / REST_THREADS;
    default: return ptiles;
    }
}

__global__ __launch_bounds__(REST_THREADS) void k_rest(Batch b) {
    __shared__ int sh_ticket, sh_fail;
    __shared__ int wsum[REST_THREADS / 64];
    __shared__ PushTileLds L;
    bool table_loaded = false;
    const int tid = threadIdx.x;
    for (int blk = blockIdx.x; blk < b.total[KC_REST]; blk += gridDim.x) {
        const int s = scene_of(b, KC_REST, blk);
        const DevScene &S = b.scenes[s];
        const uint32_t t = S.ctr->tick;
        // (scalar copies: every branch on them below is uniform for the compiler too)
        const int sub = __builtin_amdgcn_readfirstlane(substeps_of(S, t));
        if (sub <= 1) continue;
        const float h = S.dt / (float)sub;
        const int n = __builtin_amdgcn_readfirstlane(S.n), cells = __builtin_amdgcn_readfirstlane(S.cols * S.rows);
        const int tiles_r = (S.rows + PUSH_TILE - 1) / PUSH_TILE;
        const int ptiles = __builtin_amdgcn_readfirstlane(((S.cols + PUSH_TILE - 1) / PUSH_TILE) * tiles_r);
        if (S.close_on && !table_loaded) {   // the pair table is shared by every scene
            for (int q = tid; q < PAIR_SAMPLES + 4; q += REST_THREADS) L.gt[q] = S.g_lerp[q];
            table_loaded = true;   // (the first tile's barriers order these stores before its reads)
        }
        int per_sub = 0;
        for (int p = 0; p < REST_PHASES; p++) per_sub += rest_items(p, n, cells, ptiles);
        const int total = (sub - 1) * per_sub;
        int *ticket_ctr = S.rest_sync;
        int *done = S.rest_sync + 1;   // [(k - 1) * REST_PHASES + p]
        // Thread 0 publishes the block's last item (release: the block's stores,
        // ordered by the barrier that ended the item, become visible with the
        // count) and draws the next ticket in one region. The ticket and the
        // wait's outcome are read back through readfirstlane: the loop's exits
        // are then uniform branches, so no lane can come back to the ticket read
        // before thread 0 has drawn (with the ticket in a vector register the
        // compiler took the exits as divergent and sent the other lanes back to
        // the loop head without thread 0's draw, where they spun on the old ticket).
        int pending = -1;
        for (;;) {
            if (tid == 0) {
                if (pending >= 0) __hip_atomic_fetch_add(done + pending, 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
                sh_ticket = __hip_atomic_fetch_add(ticket_ctr, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
            __syncthreads();
            const int ticket = __builtin_amdgcn_readfirstlane(sh_ticket);
            __syncthreads();   // every thread holds its copy before thread 0 draws again
            if (ticket >= total) break;
            // The ticket's substep, phase and item.
            const int k = 1 + ticket / per_sub;
            int item = ticket - (k - 1) * per_sub, p = 0;
            while (item >= rest_items(p, n, cells, ptiles)) item -= rest_items(p, n, cells, ptiles), p++;
            // Wait for the phase before (the previous substep's push for phase 0).
            const int prev = p > 0 ? (k - 1) * REST_PHASES + p - 1 : (k >= 2 ? (k - 2) * REST_PHASES + REST_PHASES - 1 : -1);
            // (sh_fail is written only here, between the ticket's barriers and
            // the one below, so no thread still reads the previous item's value)
            if (tid == 0) {
                int failed = 0;
                if (prev >= 0) {
                    const int want = rest_items(p > 0 ? p - 1 : REST_PHASES - 1, n, cells, ptiles);
                    int spins = 0;
                    while (__hip_atomic_load(done + prev, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) < want) {
                        __builtin_amdgcn_s_sleep(2);
                        if (++spins > REST_SPIN_LIMIT) {
                            failed = 1;
                            __hip_atomic_store(&S.ctr->rest_error, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                            break;
                        }
                    }
                }
                sh_fail = failed;
            }
            __syncthreads();
            if (__builtin_amdgcn_readfirstlane(sh_fail)) break;
            if (p == 0) {
                for (int i = item * 1024 + tid; i < min(n, item * 1024 + 1024); i += REST_THREADS) {
                    const int cell = cell_of(S, S.cxy[i]);
                    S.erank[i] = atomicAdd(&S.count[cell], 1);
                    S.ecell[i] = cell;
                }
            } else if (p == 1 || p == 3) {
                const int c0 = item * SCAN_TILE, per = SCAN_TILE / REST_THREADS;
                int v[SCAN_TILE / REST_THREADS], sum = 0;
#pragma unroll
                for (int q = 0; q < SCAN_TILE / REST_THREADS; q++) {
                    const int c = c0 + tid * per + q;
                    v[q] = c < cells ? S.count[c] : 0;
                    sum += v[q];
                }
                int tile_total;
                int run = block_scan<REST_THREADS>(sum, wsum, tile_total);
                if (p == 1) {
                    if (tid == 0) S.toff[item] = tile_total;
                } else {
                    run += S.toff[item];
#pragma unroll
                    for (int q = 0; q < SCAN_TILE / REST_THREADS; q++) {
                        const int c = c0 + tid * per + q;
                        if (c < cells) {
                            S.cstart[c] = run;
                            S.count[c] = 0;
                        }
                        run += v[q];
                    }
                }
            } else if (p == 2) {
                const int tiles = rest_items(1, n, cells, ptiles);
                int carry = 0;
                for (int q0 = 0; q0 < tiles; q0 += REST_THREADS) {
                    const int q = q0 + tid;
                    const int v = q < tiles ? S.toff[q] : 0;
                    int chunk;
                    const int pre = block_scan<REST_THREADS>(v, wsum, chunk);
                    if (q < tiles) S.toff[q] = carry + pre;
                    carry += chunk;
                }
                if (tid == 0) S.cstart[cells] = n;
                for (int q = tid; q < ptiles; q += REST_THREADS) S.tile_dense[q] = 0;
            } else if (p == 4) {
                for (int i = item * 1024 + tid; i < min(n, item * 1024 + 1024); i += REST_THREADS)
                    S.order[S.cstart[S.ecell[i]] + S.erank[i]] = i;
            } else if (p == 5) {
                const int c = item * REST_THREADS + tid;
                if (c < cells) {
                    const int s0 = S.cstart[c], s1 = S.cstart[c + 1];
                    canon_cell(S, c, s0, s1);
                    if (s1 - s0 >= CELL_MODE_MIN)   // (every writer stores the same 1)
                        S.tile_dense[(c / S.rows / PUSH_TILE) * tiles_r + (c % S.rows) / PUSH_TILE] = 1;
                }
            } else if (S.close_on && S.tile_dense[item]) {
                push_tile<ITEM_PUSH, REST_THREADS, true>(S, L, item, k, h, t, nullptr);
            } else {   // a sparse tile: its columns' runs, one thread per electron
                const int tc0 = (item / tiles_r) * PUSH_TILE, tr0 = (item % tiles_r) * PUSH_TILE;
                const int tr1 = min(tr0 + PUSH_TILE, S.rows);
                for (int c = tc0; c < min(tc0 + PUSH_TILE, S.cols); c++) {
                    const int g0 = S.cstart[c * S.rows + tr0], g1 = S.cstart[c * S.rows + tr1];
                    for (int slot = g0 + tid; slot < g1; slot += REST_THREADS) push_one(S, slot, k, h, t);
                }
            }
            __syncthreads();   // the item's stores, before thread 0 publishes it
            pending = (k - 1) * REST_PHASES + p;
        }
    }
}

// ---------------------------------------------------------------------------------------
// Deposit + smoothing (fields.rs:18-63, 91-112; maxwell.rs:280-316): per DEP_TILE^2
// tile of the smoothing region (visible cells extended by `passes`), the cells
// whose electrons reach the tile's window add their integer CIC sums to LDS int64
// accumulators; the binomial filter (passes x [1 2 1] per axis) then runs in LDS
// and the tile writes rho, Jx, Jy on its nodes. Integer sums are independent of
// order: the result is bit-deterministic and charge is conserved exactly.

constexpr int ST = DEP_TILE;    // output tile
constexpr int MAX_PASSES = 4;
constexpr int SW = ST + 2 * MAX_PASSES;   // LDS row pitch: window of the tile plus a halo of `passes`
// Threads per deposit block: 512 (two blocks per CU at up to 128 VGPRs) where
// most tiles hold many electrons, 256 (four per CU) where most take the
// electron path below and a block's latency chain is short: the 4096^2 wire
// scene's deposit 0.88 -> 0.68 ms, while the 512^2 scene's took 46 -> 67 us.
// The results do not depend on the choice.
constexpr int DEP_THREADS_DENSE = 512, DEP_THREADS_SPARSE = 256;

// x rounded to the nearest integer (ties to even) as int64: __float2ll_rn, or,
// when the caller knows |x| < 2^31 (SMALL), the same value through the 32-bit
// conversion (two instructions instead of a dozen).
template <bool SMALL>
__device__ __forceinline__ long long rint_ll(float x) {
    return SMALL ? (long long)__float2int_rn(x) : __float2ll_rn(x);
}

// A tile whose source cells hold at most this many electrons takes them one
// per thread (12 atomics each) instead of one source cell per thread: in a
// sparse scene most source cells are empty, and a thread per cell walks them
// all for a handful of electrons.
constexpr int DEP_SPARSE_MAX = 1024;
constexpr int DEP_SRC = ST + 2 * MAX_PASSES + 2;   // source columns of a tile at most

// (held to 128 VGPRs)
template <int DEP_THREADS>
__global__ __launch_bounds__(DEP_THREADS) __attribute__((amdgpu_waves_per_eu(4))) void k_deposit_smooth(Batch b,
                                                                                                   int bias) {
    __shared__ long long dep[3][SW * SW];
    __shared__ long long tmp[ST * SW];
    __shared__ long long lcur[2][ST * ST];   // the electron path's cell current of the tile's own cells
    __shared__ int col_base[DEP_SRC + 1], col_g0[DEP_SRC];
    __shared__ float wmax[DEP_THREADS / 64];
    for (int blk = blockIdx.x; blk < b.total[KC_SMOOTH]; blk += gridDim.x) {
        const int s = scene_of(b, KC_SMOOTH, blk);
        const DevScene &S = b.scenes[s];
        const int P = S.passes;
        const int W = ST + 2 * P;
        const int local = block_in_scene(b, KC_SMOOTH, s, blk);
        const int region_c = S.cols + 2 * P, region_r = S.rows + 2 * P;
        const int tiles_r = (region_r + ST - 1) / ST;
        const int R0x = (local / tiles_r) * ST, R0y = (local % tiles_r) * ST;
        // Window cell coordinates: [wc0, wc0 + W) x [wr0, wr0 + W).
        const int wc0 = R0x - 2 * P, wr0 = R0y - 2 * P;
        const int cells = S.cols * S.rows;
        const uint32_t t = S.ctr->tick + bias;
        for (int q = threadIdx.x; q < W * W; q += DEP_THREADS) {
            const int i = q / W, j = q - (q / W) * W;
            const int c = wc0 + i, r = wr0 + j;
            const bool on = c >= 0 && r >= 0 && c < S.cols && r < S.rows;
            dep[0][i * SW + j] = on ? S.static_rho[c * S.rows + r] : 0;
            dep[1][i * SW + j] = 0;
            dep[2][i * SW + j] = 0;
        }
        const int sc0 = max(wc0 - 1, 0), sc1 = min(wc0 + W, S.cols - 1);
        const int sr0 = max(wr0 - 1, 0), sr1 = min(wr0 + W, S.rows - 1);
        const int ncol = max(sc1 - sc0 + 1, 0);
        // The source columns' electron runs (rows sr0..sr1 of a column are
        // contiguous in the sorted order); their total picks the path below.
        if (threadIdx.x < 64) {
            const int q = threadIdx.x;
            int g0 = 0, m = 0;
            if (q < ncol && sr0 <= sr1) {
                g0 = S.cstart[(sc0 + q) * S.rows + sr0];
                m = S.cstart[(sc0 + q) * S.rows + sr1 + 1] - g0;
            }
            const int incl = wave_scan_incl(m);
            if (q < ncol) {
                col_g0[q] = g0;
                col_base[q + 1] = incl;
            }
            if (q == 0) col_base[0] = 0;
        }
        for (int q = threadIdx.x; q < ST * ST; q += DEP_THREADS) {
            lcur[0][q] = 0;
            lcur[1][q] = 0;
        }
        __syncthreads();   // every entry initialised before any wave adds to it
        const int total = col_base[ncol];
        float vmax = 0.0f;
        if (total <= DEP_SPARSE_MAX) {
            // One electron per thread: its column by search over the runs, its
            // row from the snapshot's cell; CIC weights as below, each nonzero
            // node's charge product and current (rounded per electron) added to
            // the window, the cell current to the tile's own cells.
            for (int j = threadIdx.x; j < total; j += DEP_THREADS) {
                int q = 0;
                for (int step = 16; step > 0; step >>= 1)
                    if (q + step < ncol && col_base[q + step] <= j) q += step;
                const int g = col_g0[q] + (j - col_base[q]);
                const int ce = sc0 + q, re = S.scell[g] - ce * S.rows;
                const float4 e4 = S.snap[g];
                const float x = e4.x, y = e4.y, vx = e4.z, vy = e4.w;
                const bool lx = x < 0.5f, ly = y < 0.5f;
                const int fu = (int)rintf((lx ? x + 0.5f : x - 0.5f) * 4194304.0f);
                const int fv = (int)rintf((ly ? y + 0.5f : y - 0.5f) * 4194304.0f);
                const int w0x = (1 << W_BITS) - fu, w0y = (1 << W_BITS) - fv;
                const int wxn[2] = {w0x, fu}, wyn[2] = {w0y, fv};
                const int nx0 = lx ? ce - 1 : ce, ny0 = ly ? re - 1 : re;   // lower corner node
#pragma unroll
                for (int i = 0; i < 2; i++) {
                    const int li = clampi(nx0 + i, 0, S.cols - 1) - wc0;
                    if (li < 0 || li >= W) continue;
#pragma unroll
                    for (int jj = 0; jj < 2; jj++) {
                        const int lj = clampi(ny0 + jj, 0, S.rows - 1) - wr0;
                        const unsigned long long qw = (unsigned long long)(unsigned)wxn[i] * (unsigned)wyn[jj];
                        if (lj < 0 || lj >= W || qw == 0) continue;
                        const int at = li * SW + lj;
                        const float wf = __fmul_rn((float)wxn[i], (float)wyn[jj]);
                        atomicAdd((unsigned long long *)&dep[0][at], as_ull(-(long long)qw));
                        atomicAdd((unsigned long long *)&dep[1][at], as_ull(__float2ll_rn(__fmul_rn(__fmul_rn(vx, wf), J_FROM_W))));
                        atomicAdd((unsigned long long *)&dep[2][at], as_ull(__float2ll_rn(__fmul_rn(__fmul_rn(vy, wf), J_FROM_W))));
                    }
                }
                const int oc = ce - (R0x - P), orow = re - (R0y - P);
                if (oc >= 0 && oc < ST && orow >= 0 && orow < ST) {
                    const long long ix = fabsf(vx) < 8.0f ? rint_ll<true>(__fmul_rn(vx, (float)J_SCALE))
                                                          : rint_ll<false>(__fmul_rn(vx, (float)J_SCALE));
                    const long long iy = fabsf(vy) < 8.0f ? rint_ll<true>(__fmul_rn(vy, (float)J_SCALE))
                                                          : rint_ll<false>(__fmul_rn(vy, (float)J_SCALE));
                    atomicAdd((unsigned long long *)&lcur[0][oc * ST + orow], as_ull(ix));
                    atomicAdd((unsigned long long *)&lcur[1][oc * ST + orow], as_ull(iy));
                }
                vmax = fmaxf(vmax, sqrtf(vx * vx + vy * vy));
            }
            __syncthreads();
            for (int q = threadIdx.x; q < ST * ST; q += DEP_THREADS) {
                const int c = R0x - P + q / ST, r = R0y - P + q % ST;
                if (c < 0 || r < 0 || c >= S.cols || r >= S.rows) continue;
                S.cur[c * S.rows + r] = lcur[0][q];
                S.cur[cells + c * S.rows + r] = lcur[1][q];
            }
        } else {
            // Electrons whose CIC corners reach the window live in cells [w0 - 1, w0 + W].
            // One thread per such source cell sums what the cell's electrons put on
            // the 3 x 3 nodes around it (CIC, fields.rs:27-40: the lower corner is
            // floor(x - 1/2), 22-bit fixed-point weights per axis) in registers, in
            // the cell's canonical electron order, then adds the nine sums to the
            // window with LDS atomics: 27 atomics per cell, not 12 per electron.
            // Charge: the exact int64 products of the weights. Current: the FP32 sum
            // of v wx wy over the cell, rounded to int64 once per node. Integer atomics
            // give the same bits in any order. A thread whose cell the tile owns also
            // sums the cell current (physics.rs:34-43) and the speed maximum.
            const int ncy = max(sr1 - sr0 + 1, 0), ncell = ncol * ncy;
            // The scene's arrays in registers: the LDS atomics would otherwise make the
            // compiler reload them from the scene table every iteration.
            const float4 *snap = S.snap;
            const int32_t *cstart = S.cstart;
            const int rows = S.rows, cols = S.cols;
            for (int q = threadIdx.x; q < ncell; q += DEP_THREADS) {
                const int ce = sc0 + q / ncy, re = sr0 + q % ncy;
                const int cell = ce * rows + re;
                const int k0 = cstart[cell], k1 = cstart[cell + 1];
                unsigned long long qn[9];
                float jxn[9], jyn[9];
#pragma unroll
                for (int n = 0; n < 9; n++) {
                    qn[n] = 0;
                    jxn[n] = 0.0f;
                    jyn[n] = 0.0f;
                }
                long long ix = 0, iy = 0;
                // Four electrons per round, their loads all in flight before the sums;
                // the padding slots past k1 add zero weights.
                for (int kb = k0; kb < k1; kb += 4) {
                    float xs[4], ys[4], vxs[4], vys[4];
#pragma unroll
                    for (int u = 0; u < 4; u++) {
                        const bool on = kb + u < k1;
                        const float4 e4 = snap[on ? kb + u : k0];
                        xs[u] = e4.x;
                        ys[u] = e4.y;
                        vxs[u] = on ? e4.z : 0.0f;
                        vys[u] = on ? e4.w : 0.0f;
                    }
#pragma unroll
                    for (int u = 0; u < 4; u++) {
                        if (kb + u >= k1) break;
                        const float x = xs[u], y = ys[u], vx = vxs[u], vy = vys[u];
                        const bool lx = x < 0.5f, ly = y < 0.5f;   // lower corner at ce - 1 / re - 1
                        const int fu = (int)rintf((lx ? x + 0.5f : x - 0.5f) * 4194304.0f);
                        const int fv = (int)rintf((ly ? y + 0.5f : y - 0.5f) * 4194304.0f);
                        const int w0x = (1 << W_BITS) - fu, w0y = (1 << W_BITS) - fv;
                        // weights on node columns ce - 1, ce, ce + 1 and rows re - 1, re, re + 1
                        const int wxn[3] = {lx ? w0x : 0, lx ? fu : w0x, lx ? 0 : fu};
                        const int wyn[3] = {ly ? w0y : 0, ly ? fv : w0y, ly ? 0 : fv};
#pragma unroll
                        for (int i = 0; i < 3; i++)
#pragma unroll
                            for (int j = 0; j < 3; j++) {
                                qn[3 * i + j] += (unsigned long long)(unsigned)wxn[i] * (unsigned)wyn[j];
                                const float wf = __fmul_rn((float)wxn[i], (float)wyn[j]);
                                jxn[3 * i + j] = __fmaf_rn(vx, wf, jxn[3 * i + j]);
                                jyn[3 * i + j] = __fmaf_rn(vy, wf, jyn[3 * i + j]);
                            }
                        if (fabsf(vx) < 8.0f && fabsf(vy) < 8.0f) {   // |v| 2^28 < 2^31
                            ix += rint_ll<true>(__fmul_rn(vx, (float)J_SCALE));
                            iy += rint_ll<true>(__fmul_rn(vy, (float)J_SCALE));
                        } else {
                            ix += rint_ll<false>(__fmul_rn(vx, (float)J_SCALE));
                            iy += rint_ll<false>(__fmul_rn(vy, (float)J_SCALE));
                        }
                        vmax = fmaxf(vmax, sqrtf(vx * vx + vy * vy));
                    }
                }
                // Cells owned by this tile: the cell current.
                const int oc = ce - (R0x - P), orow = re - (R0y - P);
                if (oc >= 0 && oc < ST && orow >= 0 && orow < ST) {
                    S.cur[cell] = ix;
                    S.cur[cells + cell] = iy;
                }
                if (k1 == k0) continue;
                // Nodes past the grid's edge fold onto the edge (corners clamped,
                // fields.rs:28-37); nodes outside the window belong to other tiles.
#pragma unroll
                for (int i = 0; i < 3; i++) {
                    const int li = clampi(ce - 1 + i, 0, cols - 1) - wc0;
                    if (li < 0 || li >= W) continue;
#pragma unroll
                    for (int j = 0; j < 3; j++) {
                        const int lj = clampi(re - 1 + j, 0, rows - 1) - wr0;
                        if (lj < 0 || lj >= W || qn[3 * i + j] == 0) continue;
                        const int at = li * SW + lj;
                        atomicAdd((unsigned long long *)&dep[0][at], as_ull(-(long long)qn[3 * i + j]));
                        atomicAdd((unsigned long long *)&dep[1][at], as_ull(__float2ll_rn(__fmul_rn(jxn[3 * i + j], J_FROM_W))));
                        atomicAdd((unsigned long long *)&dep[2][at], as_ull(__float2ll_rn(__fmul_rn(jyn[3 * i + j], J_FROM_W))));
                    }
                }
            }
        }
        vmax = wave_max(vmax);
        if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = vmax;
        __syncthreads();
        if (threadIdx.x == 0) {
            float m = wmax[0];
            for (int w = 1; w < DEP_THREADS / 64; w++) m = fmaxf(m, wmax[w]);
            // Every block's maximum goes to one word: atomics on one address queue
            // at its L2 channel, so a block adds its own only when it is larger
            // than the word's value (read at device scope, past the L1). The word
            // only grows within a tick, so a block that skips has nothing to add.
            unsigned *slot = &S.ctr->vmax_bits[(t + 1) & 1];
            if (__float_as_uint(m) > __hip_atomic_load(slot, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
                atomicMax(slot, __float_as_uint(m));
        }
        if (S.dep_dump) {
            for (int q = threadIdx.x; q < ST * ST; q += DEP_THREADS) {
                const int c = R0x - P + q / ST, r = R0y - P + (q % ST);
                if (c < 0 || r < 0 || c >= S.cols || r >= S.rows) continue;
                const int at = (q / ST + P) * SW + (q % ST + P);
                const int cell = c * S.rows + r;
                S.dep_dump[cell] = dep[0][at] - S.static_rho[cell];
                S.dep_dump[cells + cell] = dep[1][at];
                S.dep_dump[2 * cells + cell] = dep[2][at];
            }
        }
        // Binomial smoothing: passes of [1 2 1] = the order-2P binomial per axis.
        long long bin[2 * MAX_PASSES + 1];
        bin[0] = 1;
        for (int q = 1; q <= 2 * P; q++) {
            bin[q] = 0;
            for (int z = q; z > 0; z--) bin[z] += bin[z - 1];
        }
        // Headroom: a pass of weight 4^P stays inside int64 while every input it
        // sums is at most (2^63 - 1) / 4^P; the value at the centre of each output's
        // stencil is checked (every value some stencil reads is the centre of
        // another's). The x pass always runs in int64; the y pass does for P <= 2
        // and in FP64 beyond (its FP32 result is all that leaves the tile). A node
        // past the bound raises the scene's overflow flag.
        const long long lim = (long long)(0x7fffffffffffffffULL >> (2 * P));
        bool over = false;
#pragma unroll
        for (int ch = 0; ch < 3; ch++) {
            const float scale = ch == 0 ? S.rho_from_fixed : S.j_from_fixed;
            float *out = ch == 0 ? S.rho : (ch == 1 ? S.jx : S.jy);
            __syncthreads();
            for (int q = threadIdx.x; q < ST * W; q += DEP_THREADS) {
                const int i = q / W, j = q - (q / W) * W;
                long long acc = 0;
                for (int d = 0; d <= 2 * P; d++) acc += bin[d] * dep[ch][(i + d) * SW + j];
                const long long mid = dep[ch][(i + P) * SW + j];
                over |= mid > lim || mid < -lim;
                tmp[i * SW + j] = acc;
            }
            __syncthreads();
            for (int q = threadIdx.x; q < ST * ST; q += DEP_THREADS) {
                const int i = q / ST, j = q % ST;
                const int a = R0x + i, rr = R0y + j;
                if (a >= region_c || rr >= region_r) continue;
                const size_t node = (size_t)(a - P + S.margin) * S.ny + (rr - P + S.margin);
                if (P <= 2) {
                    long long acc = 0;
                    for (int d = 0; d <= 2 * P; d++) acc += bin[d] * tmp[i * SW + j + d];
                    const long long mid = tmp[i * SW + j + P];
                    over |= mid > lim || mid < -lim;
                    out[node] = (float)acc * scale;
                } else {
                    double acc = 0.0;
                    for (int d = 0; d <= 2 * P; d++) acc += (double)bin[d] * (double)tmp[i * SW + j + d];
                    out[node] = (float)acc * scale;
                }
            }
        }
        if (over) __hip_atomic_store(&S.ctr->charge_overflow, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------------------
// FDTD leapfrog with the absorbing margin, fused (maxwell.rs:235-270;
// fields.wgsl:106-186): Bz' = Bz - h/2 curl E, E' from curl Bz' and the edge
// currents, Bz'' = Bz' - h/2 curl E', then the damping factors. On the scene's
// last FDTD substep the kernel also writes the right-hand sides of the Gauss
// solve: -4 pi k rho and div E (maxwell.rs:170-173, 200-204; fields.wgsl:192-203).
//
// Streaming: a wave owns FDTD_STRIP corner columns (lane l at column
// j = strip base + l - 2: two halo lanes on each side) and walks fdtd_rows()
// corner rows t, reading each row of Ex, Ey, Bz, Jx, Jy once, coalesced, and writing
// each row of the results once. Column neighbours come from the next lane; the
// row before is held in registers. At step t it forms Bz'(t), Ex'(t), Ey'(t - 1)
// and Bz''(t - 1) (each needs the row after it), so it starts two rows early.
// Values outside the grid are zero, as the reference's boundaries have them;
// the arithmetic is the reference's, term for term.

// Device arrays in the global address space: accesses compile to global_*
// instructions (a generic pointer gives flat_*, whose completion count also
// covers LDS traffic such as the lane shuffles) and cannot alias the scene table.
typedef __attribute__((address_space(1))) float gfloat;

// The value of the previous / next lane (column j - 1 / j + 1) by a whole-wave
// DPP shift (wave_shr:1 = 0x138, wave_shl:1 = 0x130): a VALU move, no LDS
// traffic. The end lanes get 0; they are halo lanes whose results are not kept.
__device__ __forceinline__ float lane_prev(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x138, 0xf, 0xf, true));
}
__device__ __forceinline__ float lane_next(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x130, 0xf, 0xf, true));
}

// One corner row's loads: Ex, Bz of row t, Ey, Jx, Jy of node row t and rho of
// node row t - 2 (the right-hand side written at that step).
struct FdtdRow {
    float ex, ey, bz, jx, jy, rho;
};

__global__ __launch_bounds__(64 * FDTD_WAVES) void k_fdtd(Batch b, int f) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int blk = blockIdx.x; blk < b.total[KC_FDTD]; blk += gridDim.x) {
        const int s = scene_of(b, KC_FDTD, blk);
        const DevScene &S = b.scenes[s];
        if (f >= S.fdtd_substeps) continue;
        const int nx = S.nx, ny = S.ny, pny = S.pny;
        const int local = block_in_scene(b, KC_FDTD, s, blk);
        const int strips = (ny + 1 + FDTD_STRIP - 1) / FDTD_STRIP;
        const int sblocks = (strips + FDTD_WAVES - 1) / FDTD_WAVES;
        const int rows = fdtd_rows(nx, ny);
        const int strip = (local % sblocks) * FDTD_WAVES + wave;
        const int row_lo = (local / sblocks) * rows, row_hi = min(row_lo + rows, nx + 1);
        const int steps = row_hi - row_lo + 3;   // rows t = row_lo - 2 .. row_hi (row_lo <= nx: at least 4)
        const int j = strip * FDTD_STRIP + lane - 2;
        const bool own_col = lane >= 2 && lane < 2 + FDTD_STRIP && strip < strips && j <= ny;
        // Neighbouring waves read rows and columns this one writes: substep f
        // reads one buffer of the pair and writes the other.
        const bool odd = (f & 1) != 0;
        // (no two of these arrays overlap: the reads of later rows may move ahead
        // of the writes of earlier ones)
        const gfloat *__restrict__ src_ex = (const gfloat *)(odd ? S.ex2 : S.ex);
        const gfloat *__restrict__ src_ey = (const gfloat *)(odd ? S.ey2 : S.ey);
        const gfloat *__restrict__ src_bz = (const gfloat *)(odd ? S.bz2 : S.bz);
        gfloat *__restrict__ dst_ex = (gfloat *)(odd ? S.ex : S.ex2);
        gfloat *__restrict__ dst_ey = (gfloat *)(odd ? S.ey : S.ey2);
        gfloat *__restrict__ dst_bz = (gfloat *)(odd ? S.bz : S.bz2);
        const gfloat *__restrict__ jx = (const gfloat *)S.jx;
        const gfloat *__restrict__ jy = (const gfloat *)S.jy;
        const gfloat *__restrict__ rho = (const gfloat *)S.rho;
        gfloat *__restrict__ w1 = (gfloat *)S.w1;
        const gfloat *__restrict__ dex_x = (const gfloat *)S.damp_ex_x;
        const gfloat *__restrict__ dey_x = (const gfloat *)S.damp_ey_x;
        const gfloat *__restrict__ dbz_x = (const gfloat *)S.damp_bz_x;
        const bool rhs = f == S.fdtd_substeps - 1;
        const float hh = S.half_h, c2h = S.c2h, js = S.jscale, four_pi_k = S.four_pi_k;
        const bool jin_ex = j >= 0 && j < ny, jin_c = j >= 0 && j <= ny;   // Ex / node columns, Ey / corner columns
        const float dby = jin_c ? S.damp_bz_y[j] : 0.0f, dey = jin_c ? S.damp_ey_y[j] : 0.0f;
        const float dex = jin_ex ? S.damp_ex_y[j] : 0.0f;
        float ey_prev = 0.0f, jx_prev = 0.0f, jy_prev = 0.0f, bzp_prev = 0.0f;
        float exp_prev = 0.0f, eyp_prev = 0.0f, exd_prev = 0.0f, eyd_prev = 0.0f;
        // Loads are unconditional, at indices clamped into the arrays, and the
        // values outside the grid are zeroed afterwards; the loads of row
        // t + FDTD_AHEAD are issued before the arithmetic of row t.
        const int jc_ex = min(max(j, 0), ny - 1), jc_c = min(max(j, 0), ny);
        const int nj = j - 1, njc = min(max(nj, 0), ny - 1);
        auto load = [&](int t) {
            FdtdRow r;
            const size_t tc = (size_t)min(max(t, 0), nx), te = (size_t)min(max(t, 0), nx - 1);
            r.ex = src_ex[tc * ny + jc_ex];
            r.ey = src_ey[te * (ny + 1) + jc_c];
            r.bz = src_bz[tc * (ny + 1) + jc_c];
            r.jx = jx[te * ny + jc_ex];
            r.jy = jy[te * ny + jc_ex];
            r.rho = rhs ? rho[(size_t)min(max(t - 2, 0), nx - 1) * ny + njc] : 0.0f;
            return r;
        };
        FdtdRow ring[FDTD_AHEAD];
#pragma unroll
        for (int k = 0; k < FDTD_AHEAD; k++) ring[k] = load(row_lo - 2 + k);
        for (int base = 0; base < steps; base += FDTD_AHEAD) {
#pragma unroll
            for (int k = 0; k < FDTD_AHEAD; k++) {
                const int t = row_lo - 2 + base + k;   // the row read at this step
                const FdtdRow v = ring[k];
                if (base + k + FDTD_AHEAD < steps) ring[k] = load(t + FDTD_AHEAD);
                // (steps past the last one leave every store out: their tp is row_hi or more)
                const bool c_row = t >= 0 && t <= nx, e_row = t >= 0 && t < nx;
                const float ex_f = c_row && jin_ex ? v.ex : 0.0f;
                const float ey_f = e_row && jin_c ? v.ey : 0.0f;
                const float bz_f = c_row && jin_c ? v.bz : 0.0f;
                const float jx_f = e_row && jin_ex ? v.jx : 0.0f;
                const float jy_f = e_row && jin_ex ? v.jy : 0.0f;
                // Bz'(t, j) (maxwell.rs:207-227)
                const float curl = ((ey_f - ey_prev) - ex_f) + lane_prev(ex_f);
                const float bzp = c_row && jin_c ? bz_f + (-hh) * curl : 0.0f;
                // Ex'(t, j): edges on x = 0 and x = nx carry no current (maxwell.rs:243-248, 255-258)
                // (every lane shuffles, outside any condition: a lane left out of a
                // shuffle gives its neighbour nothing)
                const float bzp_next = lane_next(bzp), jy_prev_m = lane_prev(jy_prev);
                const float jxe = (t == 0 || t == nx) ? 0.0f : js * (jx_prev + jx_f);
                const float exp_f = c_row && jin_ex ? ex_f + (c2h * (bzp_next - bzp) - jxe) : 0.0f;
                // Ey'(t - 1, j) (maxwell.rs:259-264)
                const int tp = t - 1;
                const bool ep_row = tp >= 0 && tp < nx, cp_row = tp >= 0 && tp <= nx;
                const float jye = (j == 0 || j == ny) ? 0.0f : js * (jy_prev_m + jy_prev);
                const float eyp = ep_row && jin_c ? ey_prev - (c2h * (bzp - bzp_prev) + jye) : 0.0f;
                // Bz''(t - 1, j) and the damped fields of row t - 1
                const float curl2 = ((eyp - eyp_prev) - exp_prev) + lane_prev(exp_prev);
                const float dexx = dex_x[min(max(tp, 0), nx)], deyx = dey_x[min(max(tp, 0), nx - 1)];
                const float dbzx = dbz_x[min(max(tp, 0), nx)];
                const float exd = cp_row && jin_ex ? exp_prev * fminf(dexx, dex) : 0.0f;
                const float eyd = ep_row && jin_c ? eyp * fminf(deyx, dey) : 0.0f;
                if (own_col && tp >= row_lo && tp < row_hi) {
                    const float bzpp = bzp_prev + (-hh) * curl2;
                    dst_bz[(size_t)tp * (ny + 1) + j] = bzpp * fminf(dbzx, dby);
                    if (j < ny) dst_ex[(size_t)tp * ny + j] = exd;
                    if (tp < nx) dst_ey[(size_t)tp * (ny + 1) + j] = eyd;
                }
                // div E' at node (t - 2, j - 1), with the Gauss right-hand sides
                if (rhs) {   // uniform: every lane shuffles
                    const float div = (lane_prev(exd) - lane_prev(exd_prev)) + (eyd_prev - lane_prev(eyd_prev));
                    const int ni = t - 2;
                    if (own_col && tp >= row_lo && tp < row_hi && ni >= 0 && ni < nx && nj >= 0 && nj < ny) {
                        const size_t node = (size_t)ni * pny + nj;   // the Poisson arrays' padded rows
                        const float src = -four_pi_k * v.rho;
                        w1[node] = src;
                        w1[(size_t)nx * pny + node] = div + src;
                    }
                }
                ey_prev = ey_f;
                jx_prev = jx_f;
                jy_prev = jy_f;
                bzp_prev = bzp;
                exp_prev = exp_f;
                eyp_prev = eyp;
                exd_prev = exd;
                eyd_prev = eyd;
            }
        }
    }
}

// Right-hand sides without an FDTD step (scene creation and upload).
__global__ __launch_bounds__(256) void k_rhs(Batch b) {
    for (int blk = blockIdx.x; blk < b.total[KC_NODES]; blk += gridDim.x) {
        const int s = scene_of(b, KC_NODES, blk);
        const DevScene &S = b.scenes[s];
        const int q = block_in_scene(b, KC_NODES, s, blk) * 256 + (int)threadIdx.x;
        const int nx = S.nx, ny = S.ny;
        if (q >= nx * ny) continue;
        const int i = q / ny, j = q - (q / ny) * ny;
        const float div = (S.ex[(size_t)(i + 1) * ny + j] - S.ex[(size_t)i * ny + j]) +
                          (S.ey[(size_t)i * (ny + 1) + j + 1] - S.ey[(size_t)i * (ny + 1) + j]);
        const float src = -S.four_pi_k * S.rho[q];
        const size_t node = (size_t)i * S.pny + j;   // the Poisson arrays' padded rows
        S.w1[node] = src;
        S.w1[(size_t)nx * S.pny + node] = div + src;
    }
}

// E -= grad(phi_potential + phi_correction), phi zero beyond the grid
// (maxwell.rs:176-193, 200-204; fields.wgsl:206-237). The Poisson solve is
// linear, so the sum is solved as one potential: the second right-hand side is
// div E - 4 pi k rho (the Gauss residual), not div E. The two potentials of the
// reference nearly cancel; solved apart in FP32, each one's rounding (relative
// to the full potential) would stay in their small sum, while the residual's
// solve carries only one rounding of its right-hand side. The first channel,
// P(-4 pi k rho), is the stored potential. With `tick` set, the scene's tick
// counter advances here, after everything that reads it.
// End of a tick for one cell (finish_tick, fields.wgsl:270-281): the displayed
// E and B at the cell centre and the smoothed (ammeter) cell current move
// towards their instantaneous values by alpha. sync = 1 (finish_sync,
// fields.wgsl:284-290, after a scene is set up or edited): the display jumps
// to the exact fields and the current average is left alone.
__device__ __forceinline__ void finish_cell(const DevScene &S, int cell, float ex, float ey, float bz, int sync) {
    const int cells = S.cols * S.rows;
    const float a = sync ? 1.0f : S.alpha_display;
    float *d = S.display;
    d[cell] += a * (ex - d[cell]);
    d[cells + cell] += a * (ey - d[cells + cell]);
    d[2 * cells + cell] += a * (bz - d[2 * cells + cell]);
    if (!sync) {
        for (int k = 0; k < 2; k++) {
            const float now = (float)S.cur[k * cells + cell] * S.i_from_fixed;
            S.cur_smooth[k * cells + cell] += S.alpha_current * (now - S.cur_smooth[k * cells + cell]);
        }
    }
}

//
// In a tick, a scene whose fields came from the second buffer (an odd number
// of FDTD substeps, so the correction's sources are not overwritten by it)
// also takes the end of its tick here (finish_cell): a node thread inside the
// visible cells forms the corrected E of its cell's far edges, (i + 1, j) and
// (i, j + 1), as the threads owning them do, and reads Bz from the buffer it
// is copied from, so the averages are those k_finish would form after this
// kernel, without a launch and a second read of the fields.
__global__ __launch_bounds__(256) void k_correct(Batch b, int tick) {
    for (int blk = blockIdx.x; blk < b.total[KC_NODES]; blk += gridDim.x) {
        const int s = scene_of(b, KC_NODES, blk);
        const DevScene &S = b.scenes[s];
        const int local = block_in_scene(b, KC_NODES, s, blk);
        const int q = local * 256 + (int)threadIdx.x;
        const int nx = S.nx, ny = S.ny;
        if (tick && local == 0 && threadIdx.x == 0) S.ctr->tick += 1;
        // After an odd number of FDTD substeps the fields are in the second buffer.
        const bool alt = tick && (S.fdtd_substeps & 1);
        const float *ex = alt ? S.ex2 : S.ex, *ey = alt ? S.ey2 : S.ey;
        if (q >= (nx + 1) * (ny + 1)) continue;
        const int i = q / (ny + 1), j = q - (q / (ny + 1)) * (ny + 1);
        const size_t pnodes = (size_t)nx * S.pny;
        auto p = [&](int a, int c) { return S.phi[pnodes + (size_t)a * S.pny + c]; };
        // Corrected Ex at (a, c), c < ny, and Ey at (a, c), a < nx.
        auto cex = [&](int a, int c) {
            const float high = a < nx ? p(a, c) : 0.0f, low = a > 0 ? p(a - 1, c) : 0.0f;
            return ex[(size_t)a * ny + c] - (high - low);
        };
        auto cey = [&](int a, int c) {
            const float high = c < ny ? p(a, c) : 0.0f, low = c > 0 ? p(a, c - 1) : 0.0f;
            return ey[(size_t)a * (ny + 1) + c] - (high - low);
        };
        float exn = 0.0f, eyn = 0.0f;
        if (j < ny) S.ex[(size_t)i * ny + j] = exn = cex(i, j);
        if (i < nx) S.ey[(size_t)i * (ny + 1) + j] = eyn = cey(i, j);
        if (alt) {
            S.bz[q] = S.bz2[q];
            const int c = i - S.margin, r = j - S.margin;
            if (c >= 0 && r >= 0 && c < S.cols && r < S.rows) {
                const float *bz = S.bz2;
                const size_t np = (size_t)ny + 1;
                const float exc = 0.5f * (exn + cex(i + 1, j));
                const float eyc = 0.5f * (eyn + cey(i, j + 1));
                const float bzc = 0.25f * (bz[(size_t)i * np + j] + bz[(size_t)(i + 1) * np + j] + bz[(size_t)i * np + j + 1] +
                                           bz[(size_t)(i + 1) * np + j + 1]);
                finish_cell(S, c * S.rows + r, exc, eyc, bzc, 0);
            }
        }
    }
}

// ---------------------------------------------------------------------------------------
// FP64 checks: the Gauss-law residual max |div E - 4 pi k rho| and the field
// energy sum Ex^2 + Ey^2 + c^2 Bz^2, per block into `out` (host sums the blocks).
__global__ __launch_bounds__(256) void k_diagnostics(const DevScene *scene, double c2, double *out, int blocks) {
    const DevScene &S = *scene;
    __shared__ double sres[256], sen[256];
    const int nx = S.nx, ny = S.ny;
    const long long corners = (long long)(nx + 1) * (ny + 1);
    double res = 0.0, en = 0.0;
    for (long long q = blockIdx.x * 256ll + threadIdx.x; q < corners; q += (long long)blocks * 256) {
        const int i = (int)(q / (ny + 1)), j = (int)(q % (ny + 1));
        if (i < nx && j < ny) {
            const double div = ((double)S.ex[(size_t)(i + 1) * ny + j] - (double)S.ex[(size_t)i * ny + j]) +
                               ((double)S.ey[(size_t)i * (ny + 1) + j + 1] - (double)S.ey[(size_t)i * (ny + 1) + j]);
            const double want = (double)S.four_pi_k * (double)S.rho[(size_t)i * ny + j];
            res = fmax(res, fabs(div - want));
        }
        const double bz = S.bz[q];
        en += c2 * bz * bz;
        if (j < ny) {
            const double e = S.ex[(size_t)i * ny + j];
            en += e * e;
        }
        if (i < nx) {
            const double e = S.ey[(size_t)i * (ny + 1) + j];
            en += e * e;
        }
    }
    sres[threadIdx.x] = res;
    sen[threadIdx.x] = en;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) {
            sres[threadIdx.x] = fmax(sres[threadIdx.x], sres[threadIdx.x + o]);
            sen[threadIdx.x] += sen[threadIdx.x + o];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[2 * blockIdx.x] = sres[0];
        out[2 * blockIdx.x + 1] = sen[0];
    }
}

// ---------------------------------------------------------------------------------------
// Kernel sequences (host side).



static inline int grid_of(int total) { return total < 8192 ? total : 8192; }

#define EMB_LAUNCH(kernel, kc, threads, ...)                                         \
    do {                                                                             \
        if (b.total[kc] > 0) kernel<<<grid_of(b.total[kc]), threads, 0, stream>>>(__VA_ARGS__); \
    } while (0)

// The deposit with the block size that suits the batch: the wide blocks when
// its electrons average at least 512 per tile (KC_ELECTRONS counts blocks of 256).
static void launch_deposit(const Batch &b, hipStream_t stream, int bias) {
    if (b.total[KC_ELECTRONS] >= 2 * b.total[KC_SMOOTH])
        EMB_LAUNCH(k_deposit_smooth<DEP_THREADS_DENSE>, KC_SMOOTH, DEP_THREADS_DENSE, b, bias);
    else
        EMB_LAUNCH(k_deposit_smooth<DEP_THREADS_SPARSE>, KC_SMOOTH, DEP_THREADS_SPARSE, b, bias);
}

// Counting sort of the electrons by cell: counts, scan, scatter, canonical order.
hipError_t sort_kernels(const Batch &b, hipStream_t stream) {
    EMB_LAUNCH(k_count, KC_ELECTRONS, 256, b);
    EMB_LAUNCH(k_scan, KC_SCAN, 256, b);
    if (b.top_small) {
        EMB_LAUNCH(k_scatter<true>, KC_ELECTRONS, 256, b);
    } else {
        EMB_LAUNCH(k_scan_top, KC_SCENE, TOP_THREADS, b);
        EMB_LAUNCH(k_scatter<false>, KC_ELECTRONS, 256, b);
    }
    EMB_LAUNCH(k_canon, KC_CELLS, 256, b);
    return hipGetLastError();
}

// Sources and Gauss's law for the present electrons, with no FDTD step
// (Maxwell::for_world and GpuSim::upload: maxwell.rs:88-94, sim.rs:518-528).
hipError_t setup_kernels(const Batch &b, hipStream_t stream) {
    hipError_t e = sort_kernels(b, stream);
    if (e != hipSuccess) return e;
    launch_deposit(b, stream, -1);
    // Gauss's law enforced twice (Maxwell::enforce_gauss, maxwell.rs:176-193).
    // From fields that are far from it (a new scene starts at E = 0) the first
    // pass solves a full potential, whose FP32 rounding stays in E; the second
    // solves the small residual that rounding left, and is a no-op in exact
    // arithmetic, as on fields that already satisfy the law.
    for (int pass = 0; pass < 2; pass++) {
        EMB_LAUNCH(k_rhs, KC_NODES, 256, b);
        e = launch_dst_gemms(b, stream, DST_BOTH);
        if (e != hipSuccess) return e;
        EMB_LAUNCH(k_correct, KC_NODES, 256, b, 0);
    }
    return hipGetLastError();
}

// The end of a tick (finish_cell) at every cell of the scenes whose k_correct
// did not take it (sync, or an even number of FDTD substeps).
__global__ __launch_bounds__(256) void k_finish(Batch b, int sync) {
    for (int blk = blockIdx.x; blk < b.total[KC_CELLS]; blk += gridDim.x) {
        const int s = scene_of(b, KC_CELLS, blk);
        const DevScene &S = b.scenes[s];
        // (a tick's k_correct did this for a scene after an odd number of FDTD substeps)
        if (!sync && (S.fdtd_substeps & 1)) continue;
        const int cells = S.cols * S.rows;
        const int cell = block_in_scene(b, KC_CELLS, s, blk) * 256 + (int)threadIdx.x;
        if (cell >= cells) continue;
        const int c = cell / S.rows, r = cell - (cell / S.rows) * S.rows;
        const int i = c + S.margin, j = r + S.margin;
        const int ny = S.ny;
        const float ex = 0.5f * (S.ex[(size_t)i * ny + j] + S.ex[(size_t)(i + 1) * ny + j]);
        const float ey = 0.5f * (S.ey[(size_t)i * (ny + 1) + j] + S.ey[(size_t)i * (ny + 1) + j + 1]);
        const float bz = 0.25f * (S.bz[(size_t)i * (ny + 1) + j] + S.bz[(size_t)(i + 1) * (ny + 1) + j] +
                                  S.bz[(size_t)i * (ny + 1) + j + 1] + S.bz[(size_t)(i + 1) * (ny + 1) + j + 1]);
        finish_cell(S, cell, ex, ey, bz, sync);
    }
}

hipError_t finish_sync_kernel(const Batch &b, hipStream_t stream) {
    EMB_LAUNCH(k_finish, KC_CELLS, 256, b, 1);
    return hipGetLastError();
}

// One tick (Simulation::step, simulation.rs:63-72). With `marks`, an event is
// recorded before each phase (TICK_PHASES of them) and one after the last.
#define EMB_MARK(p) \
    do {            \
        if (marks) (void)hipEventRecord(marks[p], stream); \
    } while (0)
hipError_t tick_kernels(const Batch &b, int fdtd_substeps, hipStream_t stream, hipEvent_t *marks) {
    EMB_MARK(0);
    {   // the tiled push's blocks, then the trim's (KC_STACKS is never empty),
        // which start as the push's last blocks leave CUs idle; a batch of many
        // stacks runs its trim alone first (beside the push its blocks would
        // each hold a push block's LDS)
        const bool fuse = b.total[KC_STACKS] <= TRIM_FUSE_MAX;
        if (!fuse) EMB_LAUNCH(k_trim, KC_STACKS, PUSH_THREADS, b);
        const int trim_grid = fuse ? grid_of(b.total[KC_STACKS]) : 0;
        const int push_grid = b.total[KC_PUSH] > 0 ? grid_of(b.total[KC_PUSH]) : 0;
        if (trim_grid + push_grid > 0) k_push_tiled<<<trim_grid + push_grid, PUSH_THREADS, 0, stream>>>(b, 0, trim_grid);
    }
    EMB_LAUNCH(k_trim_gain, KC_STACKS, 64, b);
    EMB_LAUNCH(k_push_sparse<false>, KC_ELECTRONS, 256, b, 0, nullptr);
    EMB_LAUNCH(k_push_heavy<false>, KC_HEAVY, HEAVY_THREADS, b, 0, nullptr);
    EMB_MARK(1);
    if (b.total[KC_ELECTRONS] > 0) EMB_LAUNCH(k_rest, KC_REST, REST_THREADS, b);
    EMB_MARK(2);
    EMB_LAUNCH(k_count, KC_ELECTRONS, 256, b);
    EMB_MARK(3);
    EMB_LAUNCH(k_scan, KC_SCAN, 256, b);
    if (!b.top_small) EMB_LAUNCH(k_scan_top, KC_SCENE, TOP_THREADS, b);
    EMB_MARK(4);
    if (b.top_small)
        EMB_LAUNCH(k_scatter<true>, KC_ELECTRONS, 256, b);
    else
        EMB_LAUNCH(k_scatter<false>, KC_ELECTRONS, 256, b);
    EMB_MARK(5);
    EMB_LAUNCH(k_canon, KC_CELLS, 256, b);
    EMB_MARK(6);
    launch_deposit(b, stream, 0);
    EMB_MARK(7);
    for (int f = 0; f < fdtd_substeps; f++) EMB_LAUNCH(k_fdtd, KC_FDTD, 64 * FDTD_WAVES, b, f);
    EMB_MARK(8);
    hipError_t e = launch_dst_gemms(b, stream, DST_TICK);
    if (e != hipSuccess) return e;
    EMB_MARK(9);
    EMB_LAUNCH(k_correct, KC_NODES, 256, b, 1);
    EMB_MARK(10);
    if (b.even_fdtd) EMB_LAUNCH(k_finish, KC_CELLS, 256, b, 0);
    EMB_MARK(11);
    return hipGetLastError();
}

// (the battery trim runs inside the push phase: its blocks beside the tiled
// push's, its gain kernel before the kick)
extern const char *const TICK_PHASE_NAMES[] = {"push", "rest", "count", "scan", "scatter", "canon",
                                        "deposit_smooth", "fdtd", "dst_gemm", "correct", "finish"};

hipError_t deposit_kernel(const Batch &b, hipStream_t stream) {
    launch_deposit(b, stream, -1);
    return hipGetLastError();
}

hipError_t diagnostics_kernel(const DevScene *scene, double c2, double *out, int blocks, hipStream_t stream) {
    k_diagnostics<<<blocks, 256, 0, stream>>>(scene, c2, out, blocks);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------------------
// Info-panel probe (GpuSim::encode_probe, sim.rs:663-687; probe_cell and
// probe_grid, fields.wgsl:368-466), one block over the state the last tick
// left (display and current averages, the trim's readings, the last sort's
// cell list). Output words as the reference's: 0 valid, 1 electrons in the
// cell, 2 potential, 3-4 E, 5 B, 6-7 S = (E x B) / mu0, 8-9 current, 10-11
// current across the wire and its width in cells, 12 stack state, 13-14 the
// terminal potentials, 15 stack current, 16 gain, 17 mean |E| over the wire
// cells, 18 their number; then 3 x PROBE_BINS histogram counts (E, B c^2, |S|
// of the arrow cells) over log10 bins of 18 / PROBE_BINS from 1e-12.
constexpr int PROBE_BINS = 256;

__device__ __forceinline__ void probe_hist(unsigned *h, float m) {
    if (m > 0.0f) {
        const float x = (log2f(m) * 0.30103f + 12.0f) / 18.0f * (float)PROBE_BINS;
        atomicAdd(&h[(int)fminf(fmaxf(x, 0.0f), (float)(PROBE_BINS - 1))], 1u);
    }
}

__global__ __launch_bounds__(256) void k_probe(const DevScene *scene, int cell, int stride, float c2, float inv_mu0,
                                               float *out, unsigned *hist_out) {
    __shared__ unsigned hist[3 * PROBE_BINS];
    __shared__ float psum[256], pcnt[256];
    const DevScene &S = *scene;
    const int tid = threadIdx.x, cells = S.cols * S.rows;
    for (int k = tid; k < 3 * PROBE_BINS; k += 256) hist[k] = 0u;
    __syncthreads();
    // probe_grid
    const int st = max(stride, 1);
    float sum = 0.0f, cnt = 0.0f;
    for (int q = tid; q < cells; q += 256) {
        const float ex = S.display[q], ey = S.display[cells + q], b = S.display[2 * cells + q];
        const float e = sqrtf(ex * ex + ey * ey);
        if (S.info[q].flags & F_WIRE) {
            sum += e;
            cnt += 1.0f;
        }
        probe_hist(hist + PROBE_BINS, fabsf(b) * c2);
        const int c = q / S.rows, r = q - c * S.rows;
        if (c % st == st / 2 && r % st == st / 2) {
            probe_hist(hist, e);
            probe_hist(hist + 2 * PROBE_BINS, inv_mu0 * e * fabsf(b));
        }
    }
    psum[tid] = sum;
    pcnt[tid] = cnt;
    __syncthreads();
    for (int half = 128; half > 0; half >>= 1) {
        if (tid < half) {
            psum[tid] += psum[tid + half];
            pcnt[tid] += pcnt[tid + half];
        }
        __syncthreads();
    }
    for (int k = tid; k < 3 * PROBE_BINS; k += 256) hist_out[k] = hist[k];
    if (tid != 0) return;
    out[17] = psum[0] / fmaxf(pcnt[0], 1.0f);
    out[18] = pcnt[0];
    // probe_cell
    for (int k = 0; k < 17; k++) out[k] = 0.0f;
    if (cell < 0 || cell >= cells) return;
    const int c = cell / S.rows, r = cell - c * S.rows;
    out[0] = 1.0f;
    out[1] = (float)(S.cstart[cell + 1] - S.cstart[cell]);
    // potential_at (common.wgsl:104-125): node potential, ions, electrons nearby
    float pot = S.phi[(size_t)(c + S.margin) * S.pny + (r + S.margin)];
    if (S.close_on) {
        pot += S.ion_pot[cell];
        const int R = S.reach;
        float near = 0.0f;
        for (int cc = max(c - R, 0); cc <= min(c + R, S.cols - 1); cc++)
            for (int rr = max(r - R, 0); rr <= min(r + R, S.rows - 1); rr++) {
                const int nb = cc * S.rows + rr;
                const float ox = (float)(cc - c) - 0.5f + S.half, oy = (float)(rr - r) - 0.5f + S.half;
                for (int k = S.cstart[nb]; k < S.cstart[nb + 1]; k++) {
                    const float2 f = snap_pos(S, k);
                    near += square_floor(S, 0, (ox + f.x) * S.square_samples, (oy + f.y) * S.square_samples);
                }
            }
        pot += S.k_q * near;
    }
    out[2] = pot;
    const float ex = S.display[cell], ey = S.display[cells + cell], b = S.display[2 * cells + cell];
    out[3] = ex;
    out[4] = ey;
    out[5] = b;
    out[6] = inv_mu0 * ey * b;
    out[7] = -inv_mu0 * ex * b;
    const float ix = S.cur_smooth[cell], iy = S.cur_smooth[cells + cell];
    out[8] = ix;
    out[9] = iy;
    // Clamp meter: the main direction of flow, summed across the wire.
    const int along = fabsf(ix) >= fabsf(iy) ? 0 : 1;
    float total = along == 0 ? ix : iy, width = 1.0f;
    const int dc = along == 1 ? 1 : 0, dr = along == 1 ? 0 : 1;
    for (int side = -1; side <= 1; side += 2) {
        int ac = c + side * dc, ar = r + side * dr;
        while (ac >= 0 && ar >= 0 && ac < S.cols && ar < S.rows && (S.info[ac * S.rows + ar].flags & F_CONDUCTIVE)) {
            total += S.cur_smooth[(size_t)along * cells + ac * S.rows + ar];
            width += 1.0f;
            ac += side * dc;
            ar += side * dr;
        }
    }
    out[10] = total;
    out[11] = width;
    const uint32_t flags = S.info[cell].flags;
    const int stack = (flags & F_BATTERY) ? S.bats[(flags >> F_BAT_SHIFT) - 1].stack : -1;
    if (stack >= 0) {
        const StackInfo I = S.stacks[stack];
        out[12] = (I.both_wired && S.stack_read[3 * stack + 2] != 0.0f) ? 2.0f : 1.0f;
        out[13] = S.stack_read[3 * stack];
        out[14] = S.stack_read[3 * stack + 1];
        float isum = 0.0f;
        for (int k = 0; k < I.cur_n; k++) {
            const int sc = S.stack_cells[I.cur_off + k];
            isum += S.cur_smooth[sc] * I.dir_x + S.cur_smooth[cells + sc] * I.dir_y;
        }
        out[15] = isum / (float)I.cur_n;
        out[16] = S.has_gain[stack] ? S.gains[stack] : 1.0f;
    }
}

hipError_t probe_kernel(const DevScene *scene, int cell, int stride, float c2, float inv_mu0, float *out, unsigned *hist,
                        hipStream_t stream) {
    k_probe<<<1, 256, 0, stream>>>(scene, cell, stride, c2, inv_mu0, out, hist);
    return hipGetLastError();
}

// The close-range force on every electron of the batch's scenes at their
// current (sorted) positions, written by electron index.
hipError_t close_force_kernel(const Batch &b, float2 *out, hipStream_t stream) {
    EMB_LAUNCH(k_force_tiled, KC_PUSH, FORCE_THREADS, b, out);
    EMB_LAUNCH(k_push_sparse<true>, KC_ELECTRONS, 256, b, 0, out);
    EMB_LAUNCH(k_push_heavy<true>, KC_HEAVY, HEAVY_THREADS, b, 0, out);
    return hipGetLastError();
}

}  // namespace emb
