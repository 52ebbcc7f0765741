// neighbor.hip -- neighbour rebuild on the device.
//
// Replaces Domain::pbc (domain.cpp:550-620), Neighbor::bin_atoms
// (neighbor.cpp:2034-2070), granular_bin_no_newton (neigh_gran.cpp:485-640)
// and the history carry-over of FixContactHistory::pre_exchange
// (fix_contact_history.cpp:305-426).
//
// Pipeline (one rebuild):
//   k_bin       wrap periodic coordinates, Morton key of the cell, histogram
//               (one atomic per run of equal keys in a wave)
//   scan        exclusive prefix sum over the key space (cell starts)
//   k_scatter   counting-sort scatter (one atomic per run of a cell in a wave)
//   k_cellsort  order each cell by previous index, so the result is unique
//   k_permute   gather every per-atom array into Morton order; wall lists
//               (k_permute_whist: long wall-history rows, one double a thread)
//   k_neigh<kStage>   one search over the 27 cells: row counts, and each row
//               staged with the previous-list index of the history it
//               carries (looked up by partner tag in the atom's previous row)
//   scan        row starts
//   k_neigh_compact   staged rows into the list, touching partners first,
//               history moved with whole-line stores
//   (k_neigh<kCount> + k_neigh<kFill>: the same search twice, for a rebuild
//   with a row longer than the staging rows)
//
// The membership predicate is the reference's, in FP64 without contraction:
//   rsq <= ((r_i + r_j) * cdf + skin)^2
// and history is carried iff rsq < ((r_i + r_j) * cdf)^2 and the previous
// row holds a flagged entry with the same partner tag (neigh_gran.cpp:592-612).
#include "kernels.h"

#include <algorithm>
#include <stdlib.h>

namespace dem {

__device__ inline int cell_coord(double x, int d, const Grid &G)
{
    int c = (int)floor((x - G.lo[d]) * G.inv[d]);
    c = c < 0 ? 0 : c;
    c = c >= G.nc[d] ? G.nc[d] - 1 : c;
    return c;
}

// Morton key of a cell; the cells of instance b of a batched run occupy keys
// [b*nkeys_local, (b+1)*nkeys_local), so the sort groups each instance's
// particles into one contiguous range and no stencil crosses instances
__device__ inline uint32_t cell_key(int cx, int cy, int cz, const Grid &G, int inst)
{
    return (uint32_t)inst * (uint32_t)G.nkeys_local
           + (G.spread[cx] | G.spread[G.maxc + cy] | G.spread[2 * G.maxc + cz]);
}

// Runs of equal keys in consecutive lanes of a wave (the particles are in
// the cell order of the last rebuild, so most neighbours in a wave share a
// cell): the first lane of each run does the run's one atomic.  Valid lanes
// form a prefix of the wave.  Returns the first lane of this lane's run and,
// through *len, the run's length (meaningful at the first lane).
__device__ inline int wave_run(uint32_t k, bool valid, int lane, int *len)
{
    const uint32_t kp = __shfl_up(k, 1, 64);
    const bool head = valid && (lane == 0 || kp != k);
    const unsigned long long heads = __ballot(head);
    const unsigned long long vmask = __ballot(valid);
    const unsigned long long le = lane == 63 ? ~0ull : (2ull << lane) - 1;   // lanes 0..lane
    const unsigned long long upto = heads & le;
    const unsigned long long after = heads & ~le;
    const int end = after ? __builtin_ctzll(after) : 64 - __builtin_clzll(vmask | 1ull);
    *len = end - lane;
    return upto ? 63 - __builtin_clzll(upto) : 0;
}

__global__ void __launch_bounds__(kBlock) k_bin(BinArgs a)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = i < a.n;   // every lane joins the run detection
    uint32_t k = 0;
    if (valid) {
        double4 p = a.xr[i];
        int4 im = a.img[i];
        const int inst = meta_inst((int)a.wt[i].w);   // every load before the stores
        double *pc = &p.x;
        int *ic = &im.x;
        // Domain::pbc, orthogonal box
        for (int d = 0; d < 3; d++) {
            if (!a.G.periodic[d]) continue;
            const double lo = a.G.lo[d], hi = a.hi[d];
            if (pc[d] < lo) {
                pc[d] += a.prd[d];
                ic[d]--;
            }
            if (pc[d] >= hi) {
                pc[d] -= a.prd[d];
                pc[d] = pc[d] > lo ? pc[d] : lo;
                ic[d]++;
            }
        }
        k = cell_key(cell_coord(p.x, 0, a.G), cell_coord(p.y, 1, a.G), cell_coord(p.z, 2, a.G), a.G, inst);
        a.xr[i] = p;
        a.img[i] = im;
        a.key[i] = k;
    }
    const int lane = threadIdx.x & 63;
    int len;
    const int first = wave_run(k, valid, lane, &len);
    if (valid && first == lane) atomicAdd(&a.count[k], len);
}

// ---------------------------------------------------------------------------
// Exclusive prefix sum, 4096 elements per 1024-thread block, wave64 shuffles.

constexpr int kScanThreads = 1024;
constexpr int kScanPer = 4;
constexpr int kScanTile = kScanThreads * kScanPer;

__global__ void __launch_bounds__(kScanThreads) k_scan_local(int *data, int n, int *blocksum)
{
    __shared__ int wsum[kScanThreads / 64];
    const int base = blockIdx.x * kScanTile + threadIdx.x * kScanPer;
    int v[kScanPer];
    int tsum = 0;
    for (int k = 0; k < kScanPer; k++) {
        v[k] = (base + k < n) ? data[base + k] : 0;
        tsum += v[k];
    }
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    int incl = tsum;
    for (int off = 1; off < 64; off <<= 1) {
        const int t = __shfl_up(incl, off, 64);
        if (lane >= off) incl += t;
    }
    if (lane == 63) wsum[wid] = incl;
    __syncthreads();
    if (wid == 0) {
        int w = lane < kScanThreads / 64 ? wsum[lane] : 0;
        int wi = w;
        for (int off = 1; off < 64; off <<= 1) {
            const int t = __shfl_up(wi, off, 64);
            if (lane >= off) wi += t;
        }
        if (lane < kScanThreads / 64) wsum[lane] = wi - w;  // exclusive wave offsets
        if (lane == kScanThreads / 64 - 1 && blocksum) blocksum[blockIdx.x] = wi;
    }
    __syncthreads();
    int run = wsum[wid] + incl - tsum;
    for (int k = 0; k < kScanPer; k++) {
        if (base + k < n) data[base + k] = run;
        run += v[k];
    }
}

__global__ void __launch_bounds__(kScanThreads) k_scan_add(int *data, int n, const int *blockoff)
{
    const int base = blockIdx.x * kScanTile + threadIdx.x * kScanPer;
    const int off = blockoff[blockIdx.x];
    for (int k = 0; k < kScanPer; k++)
        if (base + k < n) data[base + k] += off;
}

int scan_tmp_size(int n)
{
    int total = 0;
    int m = n;
    while (m > kScanTile) {
        m = (m + kScanTile - 1) / kScanTile;
        total += m + 1;
    }
    return total + 1;
}

void launch_scan(int *data, int n, int *tmp, hipStream_t s)
{
    const int nb = (n + kScanTile - 1) / kScanTile;
    if (nb <= 1) {
        k_scan_local<<<1, kScanThreads, 0, s>>>(data, n, nullptr);
        return;
    }
    k_scan_local<<<nb, kScanThreads, 0, s>>>(data, n, tmp);
    launch_scan(tmp, nb, tmp + nb + 1, s);
    k_scan_add<<<nb, kScanThreads, 0, s>>>(data, n, tmp);
}

// ---------------------------------------------------------------------------

__global__ void __launch_bounds__(kBlock) k_scatter(const uint32_t *key, int *cursor, int *sorted, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = i < n;
    const uint32_t k = valid ? key[i] : 0u;
    const int lane = threadIdx.x & 63;
    int len;
    const int first = wave_run(k, valid, lane, &len);
    int pos = 0;
    if (valid && first == lane) pos = atomicAdd(&cursor[k], len);   // the run's slots, in lane order
    pos = __shfl(pos, first, 64);
    if (valid) sorted[pos + lane - first] = i;
}

void launch_scatter(const uint32_t *key, int *cursor, int *sorted, int n, hipStream_t s)
{
    if (n == 0) return;
    k_scatter<<<(n + kBlock - 1) / kBlock, kBlock, 0, s>>>(key, cursor, sorted, n);
}

__global__ void __launch_bounds__(kBlock) k_cellsort(const int *cstart, int *sorted, int nkeys)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nkeys) return;
    const int b = cstart[k], e = cstart[k + 1];
    for (int p = b + 1; p < e; p++) {
        const int v = sorted[p];
        int q = p - 1;
        while (q >= b && sorted[q] > v) {
            sorted[q + 1] = sorted[q];
            q--;
        }
        sorted[q + 1] = v;
    }
}

void launch_cellsort(const int *cstart, int *sorted, int nkeys, hipStream_t s)
{
    k_cellsort<<<(nkeys + kBlock - 1) / kBlock, kBlock, 0, s>>>(cstart, sorted, nkeys);
}

void launch_bin(const BinArgs &a, hipStream_t s)
{
    if (a.n == 0) return;
    k_bin<<<(a.n + kBlock - 1) / kBlock, kBlock, 0, s>>>(a);
}

// ---------------------------------------------------------------------------

// One entry's history, dnum doubles: loads of up to 8 slots are issued
// together before their stores (a plain element loop waits out one memory
// round trip per slot, since the compiler cannot move a load above a store
// to a pointer that might alias it).
__device__ inline void copy_hist(double *__restrict__ h, const double *__restrict__ oh, int dnum)
{
    constexpr int C = 8;
    for (int d0 = 0; d0 < dnum; d0 += C) {
        double v[C];
#pragma unroll
        for (int d = 0; d < C; d++)
            if (d0 + d < dnum) v[d] = oh[d0 + d];
#pragma unroll
        for (int d = 0; d < C; d++)
            if (d0 + d < dnum) h[d0 + d] = v[d];
    }
}

constexpr int kWhistInline = 8;

// Every field of the particle is loaded before any store: the stores go
// through pointers the compiler must assume may alias the sources, so a
// load placed after a store would wait for it (one memory round trip per
// field).
__global__ void __launch_bounds__(kBlock) k_permute(PermuteArgs a)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= a.n) return;
    const int o = a.sorted[p];
    const double4 x = a.xr0[o];
    const double4 vm = a.vm0[o];
    const double4 wt = a.wt0[o];
    const int tg = a.tag0[o];
    const int og = a.orig0[o];
    const int4 im = a.img0[o];
    const int gm = a.gm0[o];
    // PrimitiveWall::buildNeighList with treshold = skin
    // (fix_wall_gran.cpp:700-703, primitive_wall_definitions.h:142-148)
    uint32_t m = 0;
    for (int k = 0; k < a.P.n; k++) {
        const double dMax = x.w + a.skin;
        // a blend rather than an index, so x stays in registers
        const int d = a.P.dim[k];
        const double xd = x.x * (double)(d == 0) + x.y * (double)(d == 1) + x.z * (double)(d == 2);
        const double dist = xd - a.P.pos[k];
        const double absdist = (dist > 0.0) ? dist : -dist;
        if (absdist <= dMax) m |= 1u << k;
    }
    a.xr1[p] = x;
    a.xhold[p] = x;
    a.vm1[p] = vm;
    a.wt1[p] = wt;
    a.tag1[p] = tg;
    a.orig1[p] = og;
    a.img1[p] = im;
    a.gm1[p] = gm;
    if (a.xo0) a.xo1[p] = a.xo0[o];
    if (a.mf0) a.mf1[p] = a.mf0[o];
    if (a.br0) a.br1[p] = a.br0[o];
    if (a.rl0)
        for (int q = 0; q < 4; q++) a.rl1[4 * (size_t)p + q] = a.rl0[4 * (size_t)o + q];
    a.wmask[p] = (uint8_t)m;
    // a short wall-history row (one batch of loads) goes here; longer ones
    // in k_permute_whist
    if (a.wdn > 0 && a.wdn <= kWhistInline)
        copy_hist(a.whist1 + (size_t)p * a.wdn, a.whist0 + (size_t)o * a.wdn, a.wdn);
}

// A long per-atom wall-history row into the new order, one double per
// thread, so the stores are whole lines (a thread per particle would store
// wdn doubles at a stride of wdn * 8 bytes: 6 planes of 3 slots, 1M
// particles, 210 us in k_permute against 56 + 78 us split).
__global__ void __launch_bounds__(kBlock) k_permute_whist(const int *sorted, const double *__restrict__ h0,
                                                         double *__restrict__ h1, int n, int wdn)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)n * wdn) return;
    const int p = (int)(t / wdn);
    const int k = (int)(t - (int64_t)p * wdn);
    h1[t] = h0[(size_t)sorted[p] * wdn + k];
}

void launch_permute(const PermuteArgs &a, hipStream_t s)
{
    if (a.n == 0) return;
    k_permute<<<(a.n + kBlock - 1) / kBlock, kBlock, 0, s>>>(a);
    if (a.wdn > kWhistInline) {
        const int64_t m = (int64_t)a.n * a.wdn;
        k_permute_whist<<<(unsigned)((m + kBlock - 1) / kBlock), kBlock, 0, s>>>(a.sorted, a.whist0, a.whist1, a.n,
                                                                             a.wdn);
    }
}

// ---------------------------------------------------------------------------

// One group of W lanes per particle (a wavefront, or half of one).  Lanes
// 0..26 of the group each own one cell of the 27-cell
// stencil (dz, dy, dx ascending, as a sequential scan visits them); the
// cells' candidates are enumerated 64 at a time, tested in parallel, and the
// passing ones compacted with a ballot + popcount prefix.  The row therefore
// comes out in exactly the order a sequential scan would write it, with
// contiguous stores.  A candidate that touches (rsq < (radsum*cdf)^2) looks
// up its history in the particle's previous row by partner tag; entries that
// carry nothing are written with flag 0 and no history (a flag-0 entry has
// never been in contact, so its history is zero by definition and the
// contact pass does not read it).
//
// MODE kCount counts each row; kFill writes the rows into the list (after
// the count and the scan), copying the carried history; kStage does the
// search once, counting each row and writing its first kStageSlots entries
// to a padded staging row as (list word, index of the carried entry in the
// previous list or -1), which k_neigh_compact then moves into the list.
constexpr int kCount = 0, kFill = 1, kStage = 2;

template <int MODE, int W, int S>
__global__ void __launch_bounds__(kBlock) k_neigh(NeighArgs a)
{
    constexpr bool FILL = MODE == kFill;
    constexpr bool STAGE = MODE == kStage;
    constexpr bool OLD = FILL || STAGE;   // looks up the previous row
    static_assert(W == 8 || W == 16 || W == 32 || W == 64, "8, 16, 32 or 64 lanes per particle");
    static_assert(S == 1 || S == 2, "cells of cutneighmax or of half of it");
    constexpr int kGroups = kBlock / W;
    constexpr int kShift = W == 64 ? 6 : (W == 32 ? 5 : (W == 16 ? 4 : 3));
    constexpr int kSide = 2 * S + 1;                 // stencil cells per dimension
    constexpr int NC = kSide * kSide * kSide;        // 27 or 125
    __shared__ int s_ex[kGroups][NC + 1], s_qb[kGroups][NC], s_code[kGroups][NC];
    constexpr int kOld = 32;                            // previous-row entries kept in LDS
    constexpr int kMap = 256;                           // candidates with a cell map
    __shared__ int s_ot[kGroups][kOld];                 // partner tags of the previous row (fill)
    __shared__ uint8_t s_cell[kGroups][kMap];           // stencil cell of each candidate
    const int lane = threadIdx.x & (W - 1);            // lane within the particle's group
    const int wv = threadIdx.x >> kShift;               // group within the block
    const int gshift = (threadIdx.x & 63) & ~(W - 1);   // the group's first bit in a wave ballot
    const uint64_t gmask = W == 64 ? ~0ull : ((((1ull << W) - 1ull)) << gshift);
    const int p = (blockIdx.x * blockDim.x + threadIdx.x) >> kShift;
    if (p >= a.n) return;  // p is uniform across the group
    // recorded after a staged pass: fills only when a row outgrew its staging row
    if (FILL && a.fill_if_over && *a.maxcount <= kStageSlots) return;
    const double4 xp = a.xr[p];
    const int inst = meta_inst((int)a.wt[p].w);
    const int cc[3] = {cell_coord(xp.x, 0, a.G), cell_coord(xp.y, 1, a.G), cell_coord(xp.z, 2, a.G)};
    // position in cell units, as cell_coord measures it, and the longest
    // cutoff any partner can have: a stencil cell whose nearest point lies
    // beyond it holds no partner and is skipped whole (the rows keep their
    // order).  A partner in the cell below along d has u < cc[d] (clamping
    // only moves particles outward from edge cells), one in the cell above
    // u >= cc[d] + 1; the 1e-6-cell margin covers the rounding of u.
    const double u[3] = {(xp.x - a.G.lo[0]) * a.G.inv[0], (xp.y - a.G.lo[1]) * a.G.inv[1],
                         (xp.z - a.G.lo[2]) * a.G.inv[2]};
    const double cutmax = (xp.w + a.rmax) * a.cdf + a.skin;
    // stencil cell k (dz, dy, dx ascending): first particle, count, image code
    auto cell_info = [&](int k, int &qb, int &cnt, int &code) {
        qb = 0;
        cnt = 0;
        code = kNoShift;
        const int off[3] = {k % kSide - S, (k / kSide) % kSide - S, k / (kSide * kSide) - S};
        int c[3], sh[3];
        bool ok = true;
        for (int d = 0; d < 3; d++) {
            c[d] = cc[d] + off[d];
            sh[d] = 0;
            if (c[d] < 0) {
                if (a.G.periodic[d]) { c[d] += a.G.nc[d]; sh[d] = -1; }
                else ok = false;
            } else if (c[d] >= a.G.nc[d]) {
                if (a.G.periodic[d]) { c[d] -= a.G.nc[d]; sh[d] = 1; }
                else ok = false;
            }
        }
        // distance to the cell's nearest face along each offset dimension:
        // to the particle's own cell boundary, plus |off| - 1 whole cells
        double d2 = 0.0;
        for (int d = 0; d < 3; d++)
            if (off[d] != 0) {
                const double g = off[d] < 0 ? (u[d] - cc[d]) + (-off[d] - 1) : (cc[d] + 1 - u[d]) + (off[d] - 1);
                const double t = fmax(g - 1e-6, 0.0) * a.cw[d];
                d2 += t * t;
            }
        if (ok && d2 <= cutmax * cutmax) {
            const uint32_t key = cell_key(c[0], c[1], c[2], a.G, inst);
            qb = a.cstart[key];
            cnt = a.cstart[key + 1] - qb;
            code = (sh[0] + 1) + 3 * (sh[1] + 1) + 9 * (sh[2] + 1);
        }
    };
    auto scan = [&](int v) {   // inclusive prefix over the group
        int incl = v;
        for (int o = 1; o < W; o <<= 1) {
            const int t = __shfl_up(incl, o, W);
            if (lane >= o) incl += t;
        }
        return incl;
    };
    // cells lane, lane + W, ... (W < NC: several cells per lane), their
    // candidate counts scanned in stencil order
    int total = 0;
#pragma unroll
    for (int r = 0; r < (NC + W - 1) / W; r++) {
        const int k = lane + r * W;
        int qb = 0, cnt = 0, code = kNoShift;
        if (k < NC) cell_info(k, qb, cnt, code);
        const int incl = scan(cnt);
        if (k < NC) {
            s_ex[wv][k] = total + incl - cnt;
            s_qb[wv][k] = qb;
            s_code[wv][k] = code;
            const int ex = total + incl - cnt;
            for (int j = 0; j < cnt && ex + j < kMap; j++) s_cell[wv][ex + j] = (uint8_t)k;
        }
        total += __shfl(incl, W - 1, W);
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");

    // the particle's previous row into LDS: partner tag of every flagged
    // entry (0 = cannot match; tags are positive), its first kOld entries
    // (a longer row is searched in memory past them)
    int ob = 0, oe = 0;
    if (OLD && a.have_old) {
        const int o = a.perm[p];
        ob = a.onstart[o];
        oe = a.onstart[o + 1];
        for (int k = lane; k < min(oe - ob, kOld); k += W) {
            const uint32_t ow = a.onj[ob + k];
            s_ot[wv][k] = (ow & (kFlagBit | kBondBit)) ? a.otag[ow & kJMask] : 0;
        }
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
    const int onl = min(oe - ob, kOld);
    const int ebase = FILL ? a.nstart[p] : 0;
    // a list longer than the allocation (possible only when the rebuild runs
    // inside a captured graph, sized beforehand) is reported, not written
    if (FILL && (int64_t)a.nstart[a.n] > a.ecap) {
        if (lane == 0) atomicOr(a.err, 8);
        return;
    }
    int written = 0;
    // the fill pass writes a row in two sweeps, touching entries then the
    // others (each in scan order): the order the staged rows are compacted
    // in, so both paths give the same list
    constexpr int kSweeps = FILL ? 2 : 1;
    for (int sweep = 0; sweep < kSweeps; sweep++)
    for (int c0 = 0; c0 < total; c0 += W) {
        const int c = c0 + lane;
        bool pass = false, touch = false;
        int q = 0, qcode = kNoShift, tq = 0;
        if (c < total) {
            int lo = 0;   // the candidate's cell: the map, or the largest k with ex[k] <= c
            if (total <= kMap) {
                lo = s_cell[wv][c];
            } else {
                int hi = NC - 1;
                while (lo < hi) {
                    const int mid = (lo + hi + 1) >> 1;
                    if (s_ex[wv][mid] <= c) lo = mid;
                    else hi = mid - 1;
                }
            }
            q = s_qb[wv][lo] + (c - s_ex[wv][lo]);
            qcode = s_code[wv][lo];
            if (OLD) tq = a.tag[q];   // issued beside the position load, not after the test
            if (!(q == p && qcode == kNoShift)) {
                // the periodic image of the candidate's cell
                const double4 xq = a.xr[q];
                int sx, sy, sz;
                shift_decode(qcode, sx, sy, sz);
                const double xj0 = xq.x + sx * a.prd[0];
                const double xj1 = xq.y + sy * a.prd[1];
                const double xj2 = xq.z + sz * a.prd[2];
                const double delx = xp.x - xj0;
                const double dely = xp.y - xj1;
                const double delz = xp.z - xj2;
                const double rsq = delx * delx + dely * dely + delz * delz;
                const double radsum = (xp.w + xq.w) * a.cdf;
                const double cutsq = (radsum + a.skin) * (radsum + a.skin);
                pass = rsq <= cutsq;
                touch = rsq < radsum * radsum;
            }
        }
        const bool emit = FILL ? pass && touch == (sweep == 0) : pass;
        const uint64_t m = (__ballot(emit) & gmask) >> gshift;
        if (OLD) {
            // first flagged entry of the previous row with the partner's tag
            // (the search of neigh_gran.cpp:596-598): only the lanes of
            // touching candidates search, in the LDS copy of the row
            const bool want = emit && touch && a.have_old;
            int found = -1;
            if (want) {
                for (int k = 0; k < onl; k++)
                    if (s_ot[wv][k] == tq) {
                        found = k;
                        break;
                    }
                if (found < 0)
                    for (int oe2 = ob + kOld; oe2 < oe; oe2++) {
                        const uint32_t ow = a.onj[oe2];
                        if ((ow & (kFlagBit | kBondBit)) && a.otag[ow & kJMask] == tq) {
                            found = oe2 - ob;
                            break;
                        }
                    }
            }
            if (STAGE && pass) {   // kFlagBit marks a touching entry here
                const int k = written + __popcll(m & ((1ull << lane) - 1ull));
                const uint32_t w = (uint32_t)q | ((uint32_t)qcode << 26) | (touch ? kFlagBit : 0u);
                if (k < kStageSlots)
                    a.stage[(size_t)p * kStageSlots + k] = make_int2((int)w, found >= 0 ? ob + found : -1);
            }
            if (FILL && emit) {
                const int e = ebase + written + __popcll(m & ((1ull << lane) - 1ull));
                uint32_t w = (uint32_t)q | ((uint32_t)qcode << 26);
                if (found >= 0) {
                    copy_hist(a.hist + (size_t)e * a.dnum, a.ohist + (size_t)(ob + found) * a.dnum, a.dnum);
                    w |= a.onj[ob + found] & (kFlagBit | kBondBit);
                } else if (a.coh_off >= 0) {
                    // the cohesion model reads its slots on entries that
                    // never touched (bond creation), so they start at zero
                    double *h = a.hist + (size_t)e * a.dnum;
                    for (int k = a.coh_off; k < a.dnum; k++) h[k] = 0.0;
                }
                a.nj[e] = w;
            }
        }
        written += __popcll(m);
    }
    // the staging row's last 64-byte granule written whole (a partial
    // granule costs a read-modify-write in HBM)
    if (STAGE) {
        const int k = written + lane;
        if (k < kStageSlots && k < ((written + 7) & ~7)) a.stage[(size_t)p * kStageSlots + k] = make_int2(0, -1);
    }
    if (!FILL && lane == 0) a.count[p] = written;
}

// The staged rows into the list, once the row lengths are scanned: G lanes
// per particle, staged entry k on lane k mod G.  A row is written touching
// entries first, then the others, each in scan order (a stable partition by
// group ballots): the histories the next steps read and write then sit side
// by side instead of between entries that do not touch.  A carried entry
// takes its flag bits from the previous list.  Then the history of the
// row's touching block (all of the row with a cohesion model), contiguous
// doubles, one double per lane with consecutive lanes on consecutive doubles
// (whole-line stores): a carried entry's slots come from the previous list,
// a new entry's are zero.  Does nothing when a row outgrew its staging row
// (the fill pass writes the list then).
template <int G>
__global__ void __launch_bounds__(kBlock) k_neigh_compact(NeighArgs a)
{
    static_assert(2 * G == kStageSlots && G <= 32, "two staging slots per lane");
    constexpr int U = 4;   // doubles in flight per lane
    __shared__ int s_src[kBlock / G][kStageSlots];   // carried entry behind each slot of the row
    const int lane = threadIdx.x & (G - 1);
    const int grp = threadIdx.x / G;
    const int gshift = (threadIdx.x & 63) & ~(G - 1);
    const int p = (blockIdx.x * blockDim.x + threadIdx.x) / G;
    if (p >= a.n) return;  // uniform across the group
    if (*a.maxcount > kStageSlots) return;
    if ((int64_t)a.nstart[a.n] > a.ecap) {
        if (lane == 0) atomicOr(a.err, 8);
        return;
    }
    const int2 *row = a.stage + (size_t)p * kStageSlots;
    const int2 st0 = row[lane];   // inside the staging row, whatever cnt is
    const int e0 = a.nstart[p];
    const int cnt = a.nstart[p + 1] - e0;
    const int2 st1 = cnt > G ? row[lane + G] : make_int2(0, -1);   // rows past G entries are rare
    const bool v0 = lane < cnt, v1 = lane + G < cnt;
    const bool tc0 = v0 && ((uint32_t)st0.x & kFlagBit), tc1 = v1 && ((uint32_t)st1.x & kFlagBit);
    auto group_mask = [&](bool b) { return (uint32_t)((__ballot(b) >> gshift) & ((1ull << G) - 1ull)); };
    const uint32_t mt0 = group_mask(tc0), mt1 = group_mask(tc1);
    const uint32_t mn0 = group_mask(v0 && !tc0), mn1 = group_mask(v1 && !tc1);
    const uint32_t below = (1u << lane) - 1u;
    const int nt = __popc(mt0) + __popc(mt1);
    const int d0 = tc0 ? __popc(mt0 & below) : nt + __popc(mn0 & below);
    const int d1 = tc1 ? __popc(mt0) + __popc(mt1 & below) : nt + __popc(mn0) + __popc(mn1 & below);
    if (v0) {
        a.nj[e0 + d0] = ((uint32_t)st0.x & ~kFlagBit) | (st0.y >= 0 ? a.onj[st0.y] & (kFlagBit | kBondBit) : 0u);
        s_src[grp][d0] = st0.y;
    }
    if (v1) {
        a.nj[e0 + d1] = ((uint32_t)st1.x & ~kFlagBit) | (st1.y >= 0 ? a.onj[st1.y] & (kFlagBit | kBondBit) : 0u);
        s_src[grp][d1] = st1.y;
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const int dnum = a.dnum;
    if (dnum == 0) return;
    // the touching block (the first nt entries) takes its history; the
    // entries after it have flag 0, and an unflagged entry's history is not
    // read (a first contact starts from zero), so it is left unwritten,
    // except that the cohesion models read their slots on every entry
    const int H = (a.coh_off >= 0 ? cnt : nt) * dnum;   // <= kStageSlots * dnum < 2^11
    // t / dnum as (t * inv) >> 20: exact for t < 2^11, dnum <= 64
    const uint32_t inv = (1u << 20) / (uint32_t)dnum + 1u;
    const double *__restrict__ oh = a.ohist;
    double *__restrict__ hb = a.hist + (size_t)e0 * dnum;
    for (int t0 = 0; t0 < H; t0 += U * G) {
        double v[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int t = t0 + u * G + lane;
            const int tt = min(t, H - 1);   // lanes past the block read inside it
            const int k = (int)(((uint32_t)tt * inv) >> 20);
            const int d = tt - k * dnum;
            const int src = s_src[grp][k];
            v[u] = (t < H && src >= 0) ? oh[(size_t)src * dnum + d] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int t = t0 + u * G + lane;
            if (t < H) hb[t] = v[u];
        }
    }
}

// Longest row, for the neigh_modify one check: a grid-stride max over the
// row lengths with one atomic per block (a per-row atomic on one address
// serialises a million times per rebuild).
constexpr int kMaxBlocks = 512;

__global__ void __launch_bounds__(kBlock) k_row_max(const int *count, int n, int *maxcount)
{
    __shared__ int s_m[kBlock / 64];
    int m = 0;
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) m = max(m, count[i]);
    for (int o = 32; o > 0; o >>= 1) m = max(m, __shfl_xor(m, o, 64));
    if ((threadIdx.x & 63) == 0) s_m[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < kBlock / 64; w++) m = max(m, s_m[w]);
        atomicMax(maxcount, m);
    }
}

// lanes per particle of the neighbour build: DEM_NEIGH_LANES (8, 16, 32 or
// 64), default 16 (1M-sphere column, single search: 494 us per stage pass
// against 492 with 8 lanes and 569 with 32; two-search form: 1.86 ms per
// rebuild against 1.98 with 32 lanes and 2.34 with 64; the lists are
// identical)
static int neigh_lanes()
{
    static int w = 0;
    if (w == 0) {
        const char *s = getenv("DEM_NEIGH_LANES");
        const int v = s ? atoi(s) : 16;
        w = (v == 8 || v == 32 || v == 64) ? v : 16;
    }
    return w;
}

static int group_blocks(int n, int W)
{
    const int per_block = kBlock / W;
    return (int)(((long long)n + per_block - 1) / per_block);
}

// the search kernel for the grid's cell size (reach cells per cutneighmax)
template <int MODE, int W>
static void launch_search(const NeighArgs &a, hipStream_t s)
{
    if (a.G.reach == 2) k_neigh<MODE, W, 2><<<group_blocks(a.n, W), kBlock, 0, s>>>(a);
    else k_neigh<MODE, W, 1><<<group_blocks(a.n, W), kBlock, 0, s>>>(a);
}

void launch_neigh_count(const NeighArgs &a, hipStream_t s)
{
    if (a.n == 0) return;
    switch (neigh_lanes()) {
    case 64: launch_search<kCount, 64>(a, s); break;
    case 32: launch_search<kCount, 32>(a, s); break;
    case 8: launch_search<kCount, 8>(a, s); break;
    default: launch_search<kCount, 16>(a, s); break;
    }
    launch_row_max(a.count, a.n, a.maxcount, s);
}

void launch_neigh_stage(const NeighArgs &a, hipStream_t s)
{
    if (a.n == 0) return;
    switch (neigh_lanes()) {
    case 64: launch_search<kStage, 64>(a, s); break;
    case 32: launch_search<kStage, 32>(a, s); break;
    case 8: launch_search<kStage, 8>(a, s); break;
    default: launch_search<kStage, 16>(a, s); break;
    }
    launch_row_max(a.count, a.n, a.maxcount, s);
}

void launch_neigh_compact(const NeighArgs &a, hipStream_t s)
{
    if (a.n == 0) return;
    constexpr int G = kStageSlots / 2;
    k_neigh_compact<G><<<group_blocks(a.n, G), kBlock, 0, s>>>(a);
}

void launch_row_max(const int *count, int n, int *maxcount, hipStream_t s)
{
    if (n == 0) return;
    const int nb = (int)std::min<long long>(kMaxBlocks, ((long long)n + kBlock - 1) / kBlock);
    k_row_max<<<nb, kBlock, 0, s>>>(count, n, maxcount);
}

void launch_neigh_fill(const NeighArgs &a, hipStream_t s)
{
    if (a.n == 0) return;
    switch (neigh_lanes()) {
    case 64: launch_search<kFill, 64>(a, s); break;
    case 32: launch_search<kFill, 32>(a, s); break;
    case 8: launch_search<kFill, 8>(a, s); break;
    default: launch_search<kFill, 16>(a, s); break;
    }
}

// plain fills and copies: what a rebuild would otherwise do with
// hipMemsetAsync / hipMemcpyAsync, as kernels so that the rebuild can be
// recorded into a graph
__global__ void __launch_bounds__(kBlock) k_fill_int(int *p, int v, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void __launch_bounds__(kBlock) k_copy_int(int *dst, const int *src, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

// err |= bit if the longest row exceeds the limit (the graph form of the
// host check after a count pass)
__global__ void k_row_max_check(const int *maxcount, int limit, int *err, int bit)
{
    if (threadIdx.x == 0 && *maxcount > limit) atomicOr(err, bit);
}

void launch_row_max_check(const int *maxcount, int limit, int *err, int bit, hipStream_t s)
{
    k_row_max_check<<<1, 64, 0, s>>>(maxcount, limit, err, bit);
}

// A list longer than its buffers (graph stepping sizes them once): flag it
// (err |= bit) and empty every row, so that no kernel of the step reads
// entries that were not written.  The host reports the error after replay.
__global__ void k_cap_flag(const int *nstart, int n, int64_t ecap, int *flag, int *err, int bit)
{
    if (threadIdx.x == 0) {
        const int over = (int64_t)nstart[n] > ecap;
        *flag = over;
        if (over) atomicOr(err, bit);
    }
}

__global__ void __launch_bounds__(kBlock) k_cap_clear(int *nstart, int n, const int *flag)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i <= n && *flag) nstart[i] = 0;
}

void launch_cap_guard(int *nstart, int n, int64_t ecap, int *flag, int *err, int bit, hipStream_t s)
{
    k_cap_flag<<<1, 64, 0, s>>>(nstart, n, ecap, flag, err, bit);
    k_cap_clear<<<(n + 1 + kBlock - 1) / kBlock, kBlock, 0, s>>>(nstart, n, flag);
}

void launch_fill_int(int *p, int v, int n, hipStream_t s)
{
    if (n <= 0) return;
    k_fill_int<<<(n + kBlock - 1) / kBlock, kBlock, 0, s>>>(p, v, n);
}

void launch_copy_int(int *dst, const int *src, int n, hipStream_t s)
{
    if (n <= 0) return;
    k_copy_int<<<(n + kBlock - 1) / kBlock, kBlock, 0, s>>>(dst, src, n);
}

// Per-instance sums over the instance's contiguous particle range
// [cstart[b*nkl], cstart[(b+1)*nkl]): count, sum of z, sum of the observed
// plane's force, sum of mass, kinetic energy.  One block per instance,
// strided partial sums then a fixed tree: the same order on every call.
constexpr int kObsThreads = 256;
constexpr int kObs = 5;

__global__ void __launch_bounds__(kObsThreads) k_inst_obs(const int *cstart, int nkl, const double4 *xr,
                                                         const double4 *vm, const double *pwf, double *out)
{
    __shared__ double s_v[kObs][kObsThreads];
    const int b = blockIdx.x;
    const int p0 = cstart[(size_t)b * nkl], p1 = cstart[(size_t)(b + 1) * nkl];
    double acc[kObs] = {0.0, 0.0, 0.0, 0.0, 0.0};
    for (int p = p0 + threadIdx.x; p < p1; p += kObsThreads) {
        const double4 x = xr[p], v = vm[p];
        acc[0] += 1.0;
        acc[1] += x.z;
        acc[2] += pwf ? pwf[p] : 0.0;
        acc[3] += v.w;
        acc[4] += 0.5 * v.w * (v.x * v.x + v.y * v.y + v.z * v.z);
    }
    for (int k = 0; k < kObs; k++) s_v[k][threadIdx.x] = acc[k];
    __syncthreads();
    for (int h = kObsThreads / 2; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h)
            for (int k = 0; k < kObs; k++) s_v[k][threadIdx.x] += s_v[k][threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x < kObs) out[(size_t)b * kObs + threadIdx.x] = s_v[threadIdx.x][0];
}

void launch_inst_obs(const int *cstart, int nkl, int ninst, const double4 *xr, const double4 *vm,
                     const double *pwf, double *out, hipStream_t s)
{
    if (ninst <= 0) return;
    k_inst_obs<<<ninst, kObsThreads, 0, s>>>(cstart, nkl, xr, vm, pwf, out);
}

}  // namespace dem
