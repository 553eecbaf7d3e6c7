// Several damped-wave iterations per launch on periodic grids: the forward
// stencil of wave_diss_iteration (solver/lib/rrgsim_wave/src/wave.cu:106-193)
// and its transpose (the potential adjoint), shared by pm_kernels.hip and
// pm_adjoint_kernels.hip (each translation unit gets its own copy in an
// unnamed namespace).
#pragma once

#include <type_traits>

#include "rrg_internal.hpp"

namespace rrg {
namespace pm {
namespace {

// Bit patterns of non-negative reals, whose unsigned order is their order
// (the maxima of a checking launch are reduced with integer atomics).
template <class real> struct MaxBits;
template <> struct MaxBits<float> { using T = unsigned int; };
template <> struct MaxBits<double> { using T = unsigned long long; };

template <class real> __device__ __forceinline__ typename MaxBits<real>::T to_bits(real v) {
    if constexpr (std::is_same_v<real, float>) return __float_as_uint(v);
    else return (unsigned long long)__double_as_longlong(v);
}

template <class real> __device__ __forceinline__ real wave_max(real v) {
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
    return v;
}

// Max-reduction of a checking launch: the workgroup's maxima (waves, then
// LDS) go to one of kCheckSlots slot pairs chosen by workgroup index, so the
// atomics of a large grid spread over many addresses instead of queueing on
// one; readers take the max over the slots.  Every thread of the workgroup
// must call it.
template <class real>
__device__ __forceinline__ void check_accumulate(real dmax, real amax, typename MaxBits<real>::T* maxbits) {
    __shared__ real sd[16], sa[16];
    const int t = threadIdx.x + blockDim.x * (threadIdx.y + blockDim.y * threadIdx.z);
    const int nw = (blockDim.x * blockDim.y * blockDim.z + 63) >> 6;
    dmax = wave_max(dmax);
    amax = wave_max(amax);
    if ((t & 63) == 0) {
        sd[t >> 6] = dmax;
        sa[t >> 6] = amax;
    }
    __syncthreads();
    if (t == 0) {
        real d = sd[0], a = sa[0];
        for (int w = 1; w < nw; ++w) {
            d = fmax(d, sd[w]);
            a = fmax(a, sa[w]);
        }
        const unsigned slot = (blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z)) % unsigned(kCheckSlots);
        atomicMax(maxbits + slot, to_bits<real>(d));
        atomicMax(maxbits + kCheckSlots + slot, to_bits<real>(a));
    }
}

// K iterations per launch (K = 2, 3, 4; the adjoint at most 3) on a periodic
// NX^3 grid, NX = 64 XJ RS.
//
// Forward (ADJ false): (phi^n, phi^{n-1}, src) -> (phi^{n+K-1}, phi^{n+K}),
// every level with the arithmetic of k_wave, so one launch is bitwise K
// k_wave launches.  Adjoint (ADJ true): the transposed stencil of
// k_adj_wave (no source term) from (mu_k, mu_{k+1}) to (mu_{k-K+1},
// mu_{k-K}), with sbar accumulating each iteration's input in the same order
// as K k_adj_wave launches: sbar = ((sbar + mu_k) + mu_{k-1}) + ... (FIRST:
// the first term replaces sbar).
//
// A workgroup owns TY rows x ZB planes of full x rows and has TY + 2(K-1)
// row groups of RS waves (XJ cells per lane), rows Y0-(K-1) .. Y0+TY+K-2.
// Level 1 runs on all of them, marching z from Z0-(K-1); level L runs one
// plane behind level L-1 on the rows with a (K-L)-row halo, reading level
// L-1's y neighbours from a two-plane LDS ring and its z neighbours from
// registers (the last three planes of every level stay in registers; level
// L's "old" term is level L-2 two planes back, the input's own column for
// L = 2).
//
// x neighbours: with one wave per row (RS = 1) the wave holds the whole x
// row and takes them from registers: with XJ consecutive cells per lane
// (XC) they are the lane's own cells and one DPP wave rotate at each end, and
// rows move with wide loads and stores; with XJ chunks of 64 lanes, a DPP
// wave rotate and a lane select per chunk.  With several waves per row and
// XC (which needs YL), only the two cells beyond a wave's segment come from
// the LDS rows.  Otherwise the levels' x neighbours come from the LDS ring
// and the input's from the cache.
//
// The input's y neighbours come from the cache, or (YL) from an LDS ring of
// the input plane for all but the two edge row waves.  Loads for the next
// plane are issued PD planes ahead; the barrier only waits for LDS traffic,
// so they stay in flight across it.  HBM traffic per cell: 3 reals read and
// 2 written for K forward iterations (4K for K single launches), 3 read and
// 3 written for K adjoint iterations (5K), and one launch instead of K.
//
// Pair grids (PIN / POUT, forward FP32 with two consecutive cells per lane and
// one wave per row): consecutive launches of a stretch pass (phi^n, phi^(n-1))
// as one grid of 2 NX^3 reals in which every pair of cells (2b, 2b+1) of a row
// is stored as [phi^n(2b), phi^n(2b+1), phi^(n-1)(2b), phi^(n-1)(2b+1)], row r
// (r = y + NX z) starting at 2 NX r.  A lane then reads both levels of its two
// cells with one 16-byte load and writes the outputs with one 16-byte store,
// where the split grids take two of each.  PIN: p is such a grid (po unused);
// POUT: qa receives one (qb unused).
template <class real, int XJ> struct PlaneIn {
    real a[XJ], b[XJ], c[XJ], d[XJ], po[XJ], src[XJ], pn[XJ], sb[XJ], pq[XJ];
};

// v of lane l-1 (CTRL 0x13C, wave_ror:1) or of lane l+1 (0x134, wave_rol:1),
// rotating over the 64 lanes of the wave.
template <int CTRL> __device__ __forceinline__ float wave_rot(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xf, 0xf, false));
}
template <int CTRL> __device__ __forceinline__ double wave_rot(double v) {
    const unsigned long long u = (unsigned long long)__double_as_longlong(v);
    const unsigned lo = (unsigned)__builtin_amdgcn_update_dpp(0, int(unsigned(u)), CTRL, 0xf, 0xf, false);
    const unsigned hi = (unsigned)__builtin_amdgcn_update_dpp(0, int(unsigned(u >> 32)), CTRL, 0xf, 0xf, false);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// The x-1 and x+1 neighbours of the cells a wave holds of a periodic x row.
// XC: XJ consecutive cells per lane (x = x0 + XJ lane + j): the lane's own
// cells, and one wave rotate at each end; with one wave per row (RS = 1)
// lane 0's x-1 is lane 63's last cell (x = NX-1) and lane 63's x+1 lane 0's
// first (x = 0), with RS waves per row those two cells belong to the
// neighbouring segments and come from `row`, the whole row in LDS, at x =
// xlo and xhi.  Otherwise (RS = 1 only) XJ chunks of 64 lanes (x = 64 j +
// lane): each chunk rotated, the end lanes taking the neighbouring chunk's.
template <class real, int XJ, bool XC, int RS>
__device__ __forceinline__ void row_neighbours(const real (&r)[XJ], real (&m)[XJ], real (&p)[XJ], int lane,
                                               const real* row, unsigned xlo, unsigned xhi) {
    if constexpr (XC) {
        m[0] = wave_rot<0x13C>(r[XJ - 1]);
#pragma unroll
        for (int j = 1; j < XJ; ++j) m[j] = r[j - 1];
#pragma unroll
        for (int j = 0; j < XJ - 1; ++j) p[j] = r[j + 1];
        p[XJ - 1] = wave_rot<0x134>(r[0]);
        if constexpr (RS > 1) {
            const real lo = row[xlo], hi = row[xhi];
            if (lane == 0) m[0] = lo;
            if (lane == 63) p[XJ - 1] = hi;
        }
    } else {
        static_assert(RS == 1, "chunked rows take their x neighbours from registers with one wave per row only");
        real rr[XJ], rl[XJ];
#pragma unroll
        for (int j = 0; j < XJ; ++j) {
            rr[j] = wave_rot<0x13C>(r[j]);
            rl[j] = wave_rot<0x134>(r[j]);
        }
#pragma unroll
        for (int j = 0; j < XJ; ++j) {
            m[j] = lane == 0 ? rr[(j + XJ - 1) % XJ] : rr[j];
            p[j] = lane == 63 ? rl[(j + 1) % XJ] : rl[j];
        }
    }
}

template <class real, int K, int XJ, int RS, int TY, int ZB, int PD, bool YL, bool XC, bool ADJ, bool FIRST,
          bool OCC = false, bool PIN = false, bool POUT = false, bool CHECK = false>
__global__ __launch_bounds__(64 * RS * (TY + 2 * K - 2)) void k_wave_fused(real* __restrict__ qa,
                                                                           real* __restrict__ qb,
                                                                           const real* __restrict__ p,
                                                                           const real* __restrict__ po,
                                                                           const real* __restrict__ src,
                                                                           real* __restrict__ sbar, WaveK<real> k,
                                                                           const int* __restrict__ halt,
                                                                           const int* __restrict__ occ,
                                                                           const real* __restrict__ src_empty,
                                                                           typename MaxBits<real>::T* __restrict__ maxbits) {
    // a converged solve's remaining launches return at once (uniform, before
    // any barrier)
    if (halt != nullptr && *halt != 0) return;
    constexpr int NX = 64 * XJ * RS;
    constexpr int NY = NX / TY, NB = NY * (NX / ZB);
    constexpr int ROWS = TY + 2 * K - 2;
    constexpr int S = ZB + 2 * K - 2;  // level-1 planes Z0-(K-1) .. Z0+ZB+K-2
    static_assert(K >= 2 && (NX & (NX - 1)) == 0 && NX % TY == 0 && NX % ZB == 0, "shape");
    static_assert(RS * ROWS <= 16, "a workgroup holds at most 16 waves");
    static_assert(!XC || RS == 1 || YL, "segment ends of several waves per row come from the input plane in LDS");
    static_assert(!ADJ || K <= 3, "the adjoint's s-bar order is written for at most three levels");
    static_assert(!(PIN || POUT) || (std::is_same_v<real, float> && XJ == 2 && RS == 1 && XC && !ADJ && !OCC),
                  "pair grids: forward FP32 shapes with two consecutive cells per lane and one wave per row");
    static_assert(!CHECK || (!ADJ && !PIN && !POUT), "checking launches: forward, split grids");
    __shared__ real lq[K - 1][2][ROWS][NX];
    __shared__ real l0[YL ? 2 : 1][YL ? ROWS : 1][NX];  // (YL) the input plane, for y neighbours
    int b = blockIdx.x;
    if constexpr ((NB & 7) == 0) b = (b & 7) * (NB >> 3) + (b >> 3);  // one z slab per XCD
    const int by = b % NY, bz = b / NY;
    // the wave index is uniform: row offsets live in scalar registers and every
    // global access is a scalar base plus the lane's column offset
    const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), l = threadIdx.x & 63;
    const int w = wid / RS, seg = wid - w * RS;
    const int Y0 = by * TY, Z0 = bz * ZB;
    const int y = (Y0 - (K - 1) + w + NX) & (NX - 1);
    const int ym = (y - 1) & (NX - 1), yp = (y + 1) & (NX - 1);
    const bool out_row = w >= K - 1 && w < K - 1 + TY;  // rows Y0 .. Y0+TY-1
    // Issue priority by the levels a row computes (row w computes levels 1 ..
    // min(w, ROWS-1-w, K-1) + 1): the rows that compute every level are each
    // barrier step's critical path, so the SIMD arbiter serves them first and
    // the halo rows fill the gaps.  Graph chains on MI355X, per iteration:
    // 128^3 FP32 K=4 2.87 -> 2.77 us, FP64 4.93 -> 4.79 us, 64^3 FP32 and
    // FP64 K=3 -2% and -3%, 256^3 FP32 unchanged; with several waves per row
    // (512^3 FP32 192 -> 200 us, 256^3 FP64 +1%) it loses, so those shapes
    // keep one priority.  Results are unchanged either way.
    if constexpr (RS == 1) {
        const int lv = min(min(w, ROWS - 1 - w), K - 1);
        if (lv == K - 1) __builtin_amdgcn_s_setprio(3);
        else if (lv >= 1) __builtin_amdgcn_s_setprio(1);
    }
    const bool edge = !YL || w == 0 || w == ROWS - 1;    // loads the input's y neighbours
    // XC: XJ consecutive cells per lane; otherwise XJ chunks of 64 lanes
    unsigned x[XJ], xm[XJ], xp[XJ];
#pragma unroll
    for (int j = 0; j < XJ; ++j) {
        x[j] = XC ? unsigned(seg * 64 * XJ + XJ * l + j) : unsigned(seg * 64 * XJ + 64 * j + l);
        xm[j] = (x[j] - 1u) & unsigned(NX - 1);
        xp[j] = (x[j] + 1u) & unsigned(NX - 1);
    }
    // the cells beyond this wave's segment of the row (several waves per row)
    const unsigned xlo = unsigned(seg * 64 * XJ - 1) & unsigned(NX - 1);
    const unsigned xhi = unsigned((seg + 1) * 64 * XJ) & unsigned(NX - 1);
    // offset of row yy of plane z (z wraps)
    auto row = [](int z, int yy) { return size_t(unsigned((z + NX) & (NX - 1)) * unsigned(NX) + unsigned(yy)) * NX; };
    const int zs = Z0 - (K - 1);  // level-1 plane of step 0
    // (OCC: forward, with occ) bit t: the source row of step t holds mass.  For
    // the rows without, only lane 0 issues the row's load (of its own
    // phi^(n-1) cell: a cache hit; every path issues the same load
    // instruction, so the waits on the prefetched planes stay exact, and a
    // one-lane load costs the memory pipeline almost nothing) and the step
    // uses se, the empty-cell value, instead.  Lane t fetches the flags of step t (the
    // 64-cell segments seg*XJ .. seg*XJ+XJ-1 of its row) with one vector
    // load issued before the planes' loads.  The first PD planes read their
    // rows regardless.  (Shapes marching more than 64 steps read every row.)
    static_assert(!OCC || (!ADJ && S <= 64), "source flags: forward shapes of at most 64 steps");
    unsigned long long occm = ~0ull;
    real se = 0;
    if constexpr (OCC) {
        {
            se = *src_empty;
            int f = 0;
            if (l >= PD && l < S) {
                const int* __restrict__ fp = occ + (row(zs + l, y) >> 6) + size_t(seg) * XJ;
#pragma unroll
                for (int i = 0; i < XJ; ++i) f |= fp[i];
            }
            occm = __ballot(f != 0) | ((1ull << PD) - 1);
        }
    }
    auto load = [&](PlaneIn<real, XJ>& in, int t) {
        const int zq = zs + t;
        const real* __restrict__ r0 = p + row(zq, y);
        const real* __restrict__ rm = p + row(zq, ym);
        const real* __restrict__ rp = p + row(zq, yp);
        const real* __restrict__ r2 = p + row(zq + 2, y);
        const real* __restrict__ ro = po + row(zq, y);
        if constexpr (RS > 1 && !XC) {
#pragma unroll
            for (int j = 0; j < XJ; ++j) {
                in.a[j] = r0[xm[j]];
                in.b[j] = r0[xp[j]];
            }
        }
        if constexpr (PIN) {
            // phi^n of the y neighbours (the first half of each lane's block);
            // both levels of plane z+2 (phi^(n-1) of it is used two steps on)
            if (edge) {
                const float2 cm = *reinterpret_cast<const float2*>(p + 2 * row(zq, ym) + 4 * l);
                const float2 cp = *reinterpret_cast<const float2*>(p + 2 * row(zq, yp) + 4 * l);
                in.c[0] = cm.x;
                in.c[1] = cm.y;
                in.d[0] = cp.x;
                in.d[1] = cp.y;
            }
            const float4 q2 = *reinterpret_cast<const float4*>(p + 2 * row(zq + 2, y) + 4 * l);
            in.pn[0] = q2.x;
            in.pn[1] = q2.y;
            in.pq[0] = q2.z;
            in.pq[1] = q2.w;
        } else {
            if (edge) {
#pragma unroll
                for (int j = 0; j < XJ; ++j) {
                    in.c[j] = rm[x[j]];
                    in.d[j] = rp[x[j]];
                }
            }
#pragma unroll
            for (int j = 0; j < XJ; ++j) {
                in.po[j] = ro[x[j]];
                in.pn[j] = r2[x[j]];
            }
        }
        if constexpr (OCC) {
            // a row without mass: lane 0 alone loads (its phi^(n-1) cell), so
            // every path issues the same load instruction
            const bool full = (occm >> t) & 1ull;
            const real* __restrict__ rs = full ? src + row(zq, y) : ro;
            if (full || l == 0) {
#pragma unroll
                for (int j = 0; j < XJ; ++j) in.src[j] = rs[x[j]];
            }
        } else if constexpr (!ADJ) {
            const real* __restrict__ rs = src + row(zq, y);
#pragma unroll
            for (int j = 0; j < XJ; ++j) in.src[j] = rs[x[j]];
        } else if (!FIRST && t >= 2 * (K - 1) && out_row) {
            const real* __restrict__ rb = sbar + row(Z0 + t - 2 * (K - 1), y);
#pragma unroll
            for (int j = 0; j < XJ; ++j) in.sb[j] = rb[x[j]];
        }
    };
    // the input at planes z+1, z, z-1, z-2 of the level-1 plane z; every
    // level's value of the previous step (h1) and of two steps back (h2); the
    // source of the level-1 plane of the last K steps (level L at step s is on
    // the level-1 plane of step s-L+1)
    real Pv[4][XJ];
    real h1[K + 1][XJ], h2[K + 1][XJ], sh[K][XJ];
    real pq[2][XJ];  // (PIN) phi^(n-1) of the level-1 planes of steps s and s+1, slot s & 1
    // (CHECK) max|phi^(n+K) - phi^(n+K-1)| and max|phi^(n+K)| over the cells
    // this workgroup writes: k_wave's check of the last level
    real dmax = 0, amax = 0;
    {
        if constexpr (PIN) {
            const float2 a = *reinterpret_cast<const float2*>(p + 2 * row(zs - 1, y) + 4 * l);
            const float4 c = *reinterpret_cast<const float4*>(p + 2 * row(zs, y) + 4 * l);
            const float4 d = *reinterpret_cast<const float4*>(p + 2 * row(zs + 1, y) + 4 * l);
            Pv[2][0] = a.x;
            Pv[2][1] = a.y;
            Pv[1][0] = c.x;
            Pv[1][1] = c.y;
            pq[0][0] = c.z;
            pq[0][1] = c.w;
            Pv[0][0] = d.x;
            Pv[0][1] = d.y;
            pq[1][0] = d.z;
            pq[1][1] = d.w;
        }
        const real* __restrict__ a = p + row(zs - 1, y);
        const real* __restrict__ c = p + row(zs, y);
        const real* __restrict__ d = p + row(zs + 1, y);
#pragma unroll
        for (int j = 0; j < XJ; ++j) {
            Pv[3][j] = 0;
            if constexpr (!PIN) {
                Pv[2][j] = a[x[j]];
                Pv[1][j] = c[x[j]];
                Pv[0][j] = d[x[j]];
            }
#pragma unroll
            for (int L = 0; L <= K; ++L) h1[L][j] = h2[L][j] = 0;
#pragma unroll
            for (int i = 0; i < K; ++i) sh[i][j] = 0;
        }
    }
    PlaneIn<real, XJ> ring[PD];
#pragma unroll
    for (int i = 0; i < PD; ++i) load(ring[i], i);
    if constexpr (YL) {  // the input plane of step 0, read at step 0 from slot 1
#pragma unroll
        for (int j = 0; j < XJ; ++j) l0[1][w][x[j]] = Pv[1][j];
        __syncthreads();
    }
#pragma unroll
    for (int s = 0; s < S; ++s) {
        PlaneIn<real, XJ> in = ring[s % PD];
        if (s + PD < S) load(ring[s % PD], s + PD);
        if constexpr (PIN) {
#pragma unroll
            for (int j = 0; j < XJ; ++j) {
                in.po[j] = pq[s & 1][j];
                pq[s & 1][j] = in.pq[j];
            }
        }
        if constexpr (YL) {
            if (!edge) {
#pragma unroll
                for (int j = 0; j < XJ; ++j) {
                    in.c[j] = l0[(s - 1) & 1][w - 1][x[j]];
                    in.d[j] = l0[(s - 1) & 1][w + 1][x[j]];
                }
            }
            // the input plane of step s+1; slot s&1 was last read at step s-1
#pragma unroll
            for (int j = 0; j < XJ; ++j) l0[s & 1][w][x[j]] = Pv[0][j];
        }
        if constexpr (RS == 1 || XC) {
            const real* in_row = nullptr;  // the input plane's row w (YL), for segment ends
            if constexpr (YL) in_row = &l0[(s - 1) & 1][w][0];
            row_neighbours<real, XJ, XC, RS>(Pv[1], in.a, in.b, l, in_row, xlo, xhi);
        }
        real nw[K + 1][XJ];
#pragma unroll
        for (int j = 0; j < XJ; ++j) {
            if constexpr (!ADJ) {
                // a row without mass: the empty-cell value (the load read phi^(n-1))
                if constexpr (OCC) {
                    if (!((occm >> s) & 1ull)) in.src[j] = se;
                }
#pragma unroll
                for (int i = K - 1; i > 0; --i) sh[i][j] = sh[i - 1][j];
                sh[0][j] = in.src[j];
            }
            // level 1 (wave.cu:183-190 order; the adjoint drops the source)
            const real lap_sum = in.a[j] + in.b[j] + in.c[j] + in.d[j] + Pv[2][j] + Pv[0][j];
            if constexpr (ADJ) nw[1][j] = lap_sum * k.k_main + Pv[1][j] * k.k_actual + in.po[j] * k.k_old;
            else nw[1][j] = lap_sum * k.k_main + Pv[1][j] * k.k_actual + in.po[j] * k.k_old + in.src[j];
            lq[0][s & 1][w][x[j]] = nw[1][j];
        }
#pragma unroll
        for (int L = 2; L <= K; ++L) {
            // level L is live from step 2(L-1) on, on rows L-1 .. ROWS-L
            if (s >= 2 * (L - 1) && w >= L - 1 && w < ROWS - (L - 1)) {
                const int sl = (s - 1) & 1;
                // x neighbours of level L-1 on this row: its previous step
                real bm[XJ], bp[XJ];
                if constexpr (RS == 1 || XC) {
                    row_neighbours<real, XJ, XC, RS>(h1[L - 1], bm, bp, l, &lq[L - 2][sl][w][0], xlo, xhi);
                } else {
#pragma unroll
                    for (int j = 0; j < XJ; ++j) {
                        bm[j] = lq[L - 2][sl][w][xm[j]];
                        bp[j] = lq[L - 2][sl][w][xp[j]];
                    }
                }
#pragma unroll
                for (int j = 0; j < XJ; ++j) {
                    const real lap_sum = bm[j] + bp[j] + lq[L - 2][sl][w - 1][x[j]] + lq[L - 2][sl][w + 1][x[j]] +
                                         h2[L - 1][j] + nw[L - 1][j];
                    const real old = L == 2 ? Pv[2][j] : h2[L - 2][j];
                    if constexpr (ADJ) nw[L][j] = lap_sum * k.k_main + h1[L - 1][j] * k.k_actual + old * k.k_old;
                    else nw[L][j] = lap_sum * k.k_main + h1[L - 1][j] * k.k_actual + old * k.k_old + sh[L - 1][j];
                    if (L < K) lq[L - 1][s & 1][w][x[j]] = nw[L][j];
                }
                if (L == K && out_row) {
                    // outputs written through L2 (read by the next launch) from
                    // 128^3 on
                    const size_t o = row(Z0 + s - 2 * (K - 1), y);  // wave-uniform
                    const auto ra = through_rsrc(qa + (POUT ? 2 : 1) * o, (POUT ? 2 : 1) * NX * sizeof(real));
                    const auto rb = through_rsrc(qb + (POUT ? 0 : o), NX * sizeof(real));
                    if constexpr (CHECK) {
#pragma unroll
                        for (int j = 0; j < XJ; ++j) {
                            dmax = fmax(dmax, fabs(nw[K][j] - h1[K - 1][j]));
                            amax = fmax(amax, fabs(nw[K][j]));
                        }
                    }
                    if constexpr (POUT) {
                        const real v4[4] = {nw[K][0], nw[K][1], h1[K - 1][0], h1[K - 1][1]};
                        store_through_run<real, 4>(ra, 16u * unsigned(l), v4);
                    } else if constexpr (NX < 128) {
#pragma unroll
                        for (int j = 0; j < XJ; ++j) {
                            qa[o + x[j]] = h1[K - 1][j];
                            qb[o + x[j]] = nw[K][j];
                        }
                    } else if constexpr (XC) {
                        store_through_run<real, XJ>(ra, x[0] * unsigned(sizeof(real)), h1[K - 1]);
                        store_through_run<real, XJ>(rb, x[0] * unsigned(sizeof(real)), nw[K]);
                    } else {
#pragma unroll
                        for (int j = 0; j < XJ; ++j) {
                            store_through<real>(ra, x[j] * unsigned(sizeof(real)), h1[K - 1][j]);
                            store_through<real>(rb, x[j] * unsigned(sizeof(real)), nw[K][j]);
                        }
                    }
                    if constexpr (ADJ) {
                        // the inputs of the K iterations at this cell: the
                        // input plane K below the newest, level L (1 .. K-1)
                        // from K-L steps back
                        real acc[XJ];
#pragma unroll
                        for (int j = 0; j < XJ; ++j) {
                            acc[j] = FIRST ? Pv[K][j] : in.sb[j] + Pv[K][j];
                            if constexpr (K == 3) acc[j] = acc[j] + h2[1][j];
                            acc[j] = acc[j] + h1[K - 1][j];
                        }
                        const auto rs = through_rsrc(sbar + o, NX * sizeof(real));
                        if constexpr (NX < 128) {
#pragma unroll
                            for (int j = 0; j < XJ; ++j) sbar[o + x[j]] = acc[j];
                        } else if constexpr (XC) {
                            store_through_run<real, XJ>(rs, x[0] * unsigned(sizeof(real)), acc);
                        } else {
#pragma unroll
                            for (int j = 0; j < XJ; ++j) store_through<real>(rs, x[j] * unsigned(sizeof(real)), acc[j]);
                        }
                    }
                }
            } else {
#pragma unroll
                for (int j = 0; j < XJ; ++j) nw[L][j] = 0;
            }
        }
#pragma unroll
        for (int j = 0; j < XJ; ++j) {
#pragma unroll
            for (int L = 1; L <= K; ++L) {
                h2[L][j] = h1[L][j];
                h1[L][j] = nw[L][j];
            }
            Pv[3][j] = Pv[2][j];
            Pv[2][j] = Pv[1][j];
            Pv[1][j] = Pv[0][j];
            Pv[0][j] = in.pn[j];
        }
        // ring slot s&1 is rewritten at s+2, after every wave passed this barrier
        if (s + 1 < S) __syncthreads();
    }
    if constexpr (CHECK) check_accumulate<real>(dmax, amax, maxbits);
}

template <class real, int K, int XJ, int RS, int TY, int ZB, int PD, bool YL, bool XC, bool ADJ, bool FIRST,
          bool PIN = false, bool POUT = false, bool CHECK = false>
void launch_shape(hipStream_t s, real* qa, real* qb, const real* p, const real* po, const real* src, real* sbar,
                  const WaveK<real>& k, const int* halt, const int* occ, const real* src_empty,
                  typename MaxBits<real>::T* maxbits = nullptr) {
    constexpr int NX = 64 * XJ * RS;
    constexpr dim3 grid((NX / TY) * (NX / ZB)), block(64 * RS * (TY + 2 * K - 2));
    // source flags: the shapes that use them (FP64 from 128^3, FP32 from
    // 256^3, at most 64 planes marched); the others read every source row
    constexpr bool kOcc = !ADJ && ZB + 2 * K - 2 <= 64 && (NX >= 256 || (sizeof(real) == 8 && NX >= 128));
    if constexpr (PIN || POUT) {
        static_assert(!kOcc, "pair grids: shapes without source flags");
        k_wave_fused<real, K, XJ, RS, TY, ZB, PD, YL, XC, ADJ, FIRST, false, PIN, POUT>
            <<<grid, block, 0, s>>>(qa, qb, p, po, src, sbar, k, halt, nullptr, nullptr, nullptr);
    } else if (kOcc && occ != nullptr) {
        if constexpr (kOcc)
            k_wave_fused<real, K, XJ, RS, TY, ZB, PD, YL, XC, ADJ, FIRST, true, false, false, CHECK>
                <<<grid, block, 0, s>>>(qa, qb, p, po, src, sbar, k, halt, occ, src_empty, maxbits);
    } else {
        k_wave_fused<real, K, XJ, RS, TY, ZB, PD, YL, XC, ADJ, FIRST, false, false, false, CHECK>
            <<<grid, block, 0, s>>>(qa, qb, p, po, src, sbar, k, halt, nullptr, nullptr, maxbits);
    }
}

// Launch shapes (K, XJ, RS, TY, ZB, PD, YL, XC), from sweeps of
// graph-replayed 240-iteration chains of the forward kernel on MI355X (time
// per iteration; one iteration per launch in brackets):
//          FP32                                    FP64
//   64^3   K=3 (1,1,4,4,1,0,1)  1.39 us (2.74)     K=3 (1,1,4,4,1,0,1)  2.07 us (3.32)
//   128^3  K=4 (2,1,8,8,2,1,1)  3.20 (6.8)         K=4 (2,1,8,8,1,1,1)  5.15-5.40 (12.0)
//   256^3  K=4 (4,1,8,32,1,1,1) 18.3 (38.0)        K=3 (2,2,4,64,1,1,1) 45.5 (115.6)
//   512^3  K=3 (4,2,4,32,1,1,1) 200-202 (440)      K=2 (4,2,4,32,1,1,1) 578 (932); K=3
//                                                  shapes spill registers
// and for shorter remainders the K=3 and K=2 shapes below.  The LDS ring of
// the input plane (YL) and consecutive cells per lane (XC) pay on the larger
// planes and, together, at 128^3 FP32; at 128^3 FP32 K=3 with neither was
// 3.5-3.8 us, with both 3.4, and K=4 with both and a two-plane prefetch 3.2.
// K=5 gains nothing over K=4 (128^3 FP32 3.25 us, 256^3 FP32 19.8 us).  The
// adjoint uses the K<=3 shapes.  Other grids, and non-periodic ones, run one
// iteration per launch.
// The shapes with pair-grid variants (PIN / POUT): 128^3 FP32, four levels.
template <class real> bool fused_pairs(int nx, int bc, int levels) {
    return std::is_same_v<real, float> && bc == RRG_BC_PERIODIC && nx == 128 && levels == 4;
}

template <class real> int fused_max(int nx, int bc) {
    constexpr bool f32 = std::is_same_v<real, float>;
    if (bc != RRG_BC_PERIODIC) return 1;
    if (nx == 128 || (nx == 256 && f32)) return 4;
    if (nx == 64 || nx == 256 || (nx == 512 && f32)) return 3;
    return nx == 512 ? 2 : 1;
}

// Iterations of the next launch of an unchecked stretch of `remaining`
// iterations, at most `cap`.  Where four fit, a tail of five or six
// iterations runs as 3+2 and 3+3 (cheaper than 4+1 and 4+2).
template <class real> int fused_levels(int nx, int bc, int remaining, int cap = 4) {
    const int m = fused_max<real>(nx, bc) < cap ? fused_max<real>(nx, bc) : cap;
    if (m < 2 || remaining < 2) return 1;
    if (m == 2 || remaining == 2) return 2;
    if (m == 3 || remaining == 3 || remaining == 5 || remaining == 6) return 3;
    return 4;
}

// One launch of `levels` iterations with the shape above; the caller checks
// fused_levels(nx, periodic, levels) == levels, and fused_pairs for pin
// (p is a pair grid) and pout (qa receives one).  CHECK: a checking launch
// (split grids), whose last level's maxima go to maxbits.
template <class real, bool ADJ, bool FIRST, bool CHECK = false>
void launch_fused(hipStream_t s, int nx, int levels, real* qa, real* qb, const real* p, const real* po,
                  const real* src, real* sbar, const WaveK<real>& k, const int* halt = nullptr,
                  const int* occ = nullptr, const real* src_empty = nullptr, bool pin = false, bool pout = false,
                  typename MaxBits<real>::T* maxbits = nullptr) {
    constexpr bool f32 = std::is_same_v<real, float>;
    if constexpr (!ADJ) {
        if (levels == 4) {
            if (nx == 128) {
                if constexpr (f32) {
                    if (pin && pout)
                        launch_shape<real, 4, 2, 1, 8, 8, 2, true, true, ADJ, FIRST, true, true>(s, qa, qb, p, po, src, sbar, k, halt, occ, src_empty);
                    else if (pin)
                        launch_shape<real, 4, 2, 1, 8, 8, 2, true, true, ADJ, FIRST, true, false>(s, qa, qb, p, po, src, sbar, k, halt, occ, src_empty);
                    else if (pout)
                        launch_shape<real, 4, 2, 1, 8, 8, 2, true, true, ADJ, FIRST, false, true>(s, qa, qb, p, po, src, sbar, k, halt, occ, src_empty);
                    else
                        launch_shape<real, 4, 2, 1, 8, 8, 2, true, true, ADJ, FIRST, false, false, CHECK>(s, qa, qb, p, po, src, sbar, k, halt, occ, src_empty, maxbits);
                } else
                    launch_shape<real, 4, 2, 1, 8, 8, 1, true, true, ADJ, FIRST, false, false, CHECK>(s, qa, qb, p, po, src, sbar, k, halt, occ, src_empty, maxbits);
            } else if constexpr (f32) {
                launch_shape<real, 4, 4, 1, 8, 32, 1, true, true, ADJ, FIRST, false, false, CHECK>(s, qa, qb, p, po, src, sbar, k, halt, occ, src_empty, maxbits);
            }
            return;
        }
    }
    if (levels == 3) {
        if (nx == 64) {
            launch_shape<real, 3, 1, 1, 4, 4, 1, false, true, ADJ, FIRST, false, false, CHECK>(s, qa, qb, p, po, src, sbar, k, halt, occ, src_empty, maxbits);
        } else if (nx == 128) {
            launch_shape<real, 3, 2, 1, 8, 8, f32 ? 2 : 1, true, true, ADJ, FIRST, false, false, CHECK>(s, qa, qb, p, po, src, sbar, k, halt, occ, src_empty, maxbits);
        } else if (nx == 256) {
            if constexpr (f32)
                launch_shape<real, 3, 4, 1, 8, 32, 1, true, true, ADJ, FIRST, false, false, CHECK>(s, qa, qb, p, po, src, sbar, k, halt, occ, src_empty, maxbits);
            else
                launch_shape<real, 3, 2, 2, 4, 64, 1, true, true, ADJ, FIRST, false, false, CHECK>(s, qa, qb, p, po, src, sbar, k, halt, occ, src_empty, maxbits);
        } else if constexpr (f32) {
            launch_shape<real, 3, 4, 2, 4, 32, 1, true, true, ADJ, FIRST, false, false, CHECK>(s, qa, qb, p, po, src, sbar, k, halt, occ, src_empty, maxbits);
        }
    } else {
        if (nx == 64) {
            if constexpr (f32)
                launch_shape<real, 2, 1, 1, 4, 4, 2, false, true, ADJ, FIRST, false, false, CHECK>(s, qa, qb, p, po, src, sbar, k, halt, occ, src_empty, maxbits);
            else
                launch_shape<real, 2, 1, 1, 4, 4, 1, true, true, ADJ, FIRST, false, false, CHECK>(s, qa, qb, p, po, src, sbar, k, halt, occ, src_empty, maxbits);
        } else if (nx == 128) {
            if constexpr (f32)
                launch_shape<real, 2, 1, 2, 4, 8, 1, false, false, ADJ, FIRST, false, false, CHECK>(s, qa, qb, p, po, src, sbar, k, halt, occ, src_empty, maxbits);
            else
                launch_shape<real, 2, 2, 1, 8, 8, 1, true, true, ADJ, FIRST, false, false, CHECK>(s, qa, qb, p, po, src, sbar, k, halt, occ, src_empty, maxbits);
        } else if (nx == 256) {
            if constexpr (f32)
                launch_shape<real, 2, 4, 1, 8, 32, 1, true, true, ADJ, FIRST, false, false, CHECK>(s, qa, qb, p, po, src, sbar, k, halt, occ, src_empty, maxbits);
            else
                launch_shape<real, 2, 2, 2, 4, 32, 1, false, false, ADJ, FIRST, false, false, CHECK>(s, qa, qb, p, po, src, sbar, k, halt, occ, src_empty, maxbits);
        } else {
            if constexpr (f32)
                launch_shape<real, 2, 8, 1, 4, 32, 1, true, true, ADJ, FIRST, false, false, CHECK>(s, qa, qb, p, po, src, sbar, k, halt, occ, src_empty, maxbits);
            else
                launch_shape<real, 2, 4, 2, 4, 32, 1, true, true, ADJ, FIRST, false, false, CHECK>(s, qa, qb, p, po, src, sbar, k, halt, occ, src_empty, maxbits);
        }
    }
}

}  // namespace
}  // namespace pm
}  // namespace rrg
