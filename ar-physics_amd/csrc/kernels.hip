// kernels.hip -- gfx950 kernels of the particle step (SPH fluid, neighbour grid, RK
// epilogue, std430 I/O).  Built with --genco into a code object that the host runtime
// loads with hipModuleLoadData and launches with hipModuleLaunchKernel.
//
// Reference semantics are cited per kernel (paths relative to the reference repository).
#include <hip/hip_runtime.h>

#include "common.h"

// Multiply-adds are fused only inside one source expression (as C's FP_CONTRACT ON), so the
// rounding of every expression is fixed by its text and not by what the optimiser sees
// around it: the walk forms specialised for one material pair, the general walks and the
// batch entries then round alike.
#pragma clang fp contract(on)

using namespace arp;

struct AosParticle {  // ArpParticle, std430 48 B
    float den;
    uint32_t mat, solid_id, pad;
    float4 pos;
    float4 vel;
};
struct AosSolid {  // ArpSolidParticle, std430 96 B
    uint32_t part_id, pad[3];
    float4 ref_pos;
    float4 stress[4];
};

// ------------------------------------------------------------------------------------
// helpers
// ------------------------------------------------------------------------------------

// floor(clamp((p-min)/dim, 0, 1) * sub) (neighbor_list.rs:95).  A NaN quotient (axis with
// dim == 0) clamps to 0 through fmaxf; an index equal to sub (p exactly at bounds.max) is
// kept in the last cell.
__device__ __forceinline__ uint32_t cell_coord(float p, float mn, float dm, uint32_t s) {
    float t = (p - mn) / dm;
    t = fminf(fmaxf(t, 0.0f), 1.0f);
    float f = floorf(t * (float)s);
    uint32_t c = (uint32_t)f;
    return c >= s ? s - 1 : c;
}

__device__ __forceinline__ bool cell_of(const GridP& g, float x, float y, float z, uint32_t& cx,
                                        uint32_t& cy, uint32_t& cz) {
    if (!(isfinite(x) && isfinite(y) && isfinite(z))) return false;
    cx = cell_coord(x, g.min[0], g.dim[0], g.sub[0]);
    cy = cell_coord(y, g.min[1], g.dim[1], g.sub[1]);
    cz = cell_coord(z, g.min[2], g.dim[2], g.sub[2]);
    return true;
}

// Wave-aggregated atomic slot reservation: lanes with the same bin share one atomicAdd.
// Every lane of the wave must call it (valid lanes reserve a slot).
__device__ __forceinline__ uint32_t agg_rank(uint32_t* counters, uint32_t bin, bool valid) {
    // the lanes of a wave are grouped by bin (a ballot per distinct bin, no memory access),
    // then every group's first lane reserves the group's slots with one atomic, all groups
    // at once: one round trip per wave, where reserving group after group waited for each
    unsigned long long active = __ballot(valid);
    const uint32_t lane = __lane_id();
    const unsigned long long lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    uint32_t leader = lane, within = 0, count = 0;
    while (active) {
        const int l = __ffsll(active) - 1;
        const uint32_t lb = __shfl(bin, l);
        const unsigned long long m = __ballot(bin == lb) & active;
        if ((m >> lane) & 1ull) {
            leader = (uint32_t)l;
            within = (uint32_t)__popcll(m & lt);
        }
        if (lane == (uint32_t)l) count = (uint32_t)__popcll(m);
        active &= ~m;
    }
    uint32_t base = 0;
    if (valid && leader == lane) base = atomicAdd(&counters[bin], count);
    base = __shfl(base, (int)leader);
    return base + within;
}

__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
    const uint32_t lane = __lane_id();
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t t = __shfl_up(v, o);
        if (lane >= (uint32_t)o) v += t;
    }
    return v;
}

// exclusive scan over a block of 256 threads; wsum is __shared__ uint32_t[4]
__device__ __forceinline__ uint32_t block_excl_scan256(uint32_t v, uint32_t* wsum, uint32_t& total) {
    uint32_t inc = wave_incl_scan(v);
    const uint32_t wid = threadIdx.x >> 6;
    if (__lane_id() == 63) wsum[wid] = inc;
    __syncthreads();
    uint32_t pre = 0;
    for (uint32_t w = 0; w < wid; w++) pre += wsum[w];
    total = wsum[0] + wsum[1] + wsum[2] + wsum[3];
    __syncthreads();
    return pre + inc - v;
}

// Quotients of the pair terms through the hardware reciprocal (v_rcp_f32, 1 ulp) instead of
// the correctly rounded division sequence
__device__ __forceinline__ float frcp(float b) { return __builtin_amdgcn_rcpf(b); }
__device__ __forceinline__ float fdiv(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }

// the seventh power by multiplication: x^7 = x^4 x^2 x
__device__ __forceinline__ float pow7(float x) {
    const float x2 = x * x, x4 = x2 * x2;
    return x4 * x2 * x;
}

// EOS (fluid.rs:45-52).  Tait: the seventh power by multiplication and the quotients through
// v_rcp_f32, a few ulp from powf and the divisions (far below the parity tolerance) at a tenth
// of their instructions: every stage epilogue evaluates it for every particle.
__device__ __forceinline__ float eos_pressure(uint32_t eq, float density, float c, float d0) {
    switch (eq) {
        case 1: return c * c;
        case 2: return c * c * (density - d0);
        case 3: return (c * c * d0 * (1.0f / 7.0f)) * (pow7(density * frcp(d0)) - 1.0f);
        default: return 0.0f;
    }
}

// p / (den*den) of a state, the per-particle half of the pressure term (fluid.rs:183-186)
__device__ __forceinline__ float pressure_q(const DMat& m, float den) {
    return eos_pressure(m.eq, den, m.c, m.target_den) * frcp(den * den);
}

// contact_potential (fluid.rs:54-67)
__device__ __forceinline__ float contact_potential(uint32_t eq, float r, float strength, float r0) {
    switch (eq) {
        case 1: return r <= r0 ? strength : 0.0f;
        case 2: return strength * (r0 - r);
        case 3: return (strength * r * (1.0f / 7.0f)) * (pow7(r0 * frcp(r)) - 1.0f);
        case 4: {
            float f = r0 / r;
            f *= f * f;
            f *= f;
            return strength * (f * f - 2.0f * f);
        }
        case 5: return (strength * r0 * (1.0f / 7.0f)) * (pow7(r * frcp(r0)) - 1.0f);
        default: return 0.0f;
    }
}

// ------------------------------------------------------------------------------------
// neighbour grid: cell-row counting sort (replaces fill_buckets / reset_buckets,
// neighbor_list.rs:37-143)
// ------------------------------------------------------------------------------------

// Cell key of every particle + slot in its (y,z) row, reserved by wave-aggregated atomics.
__device__ __forceinline__ bool nl_reuse_prev(const uint32_t* w);

// (grid-stride over blocks of 256 particles: a gated launch that has nothing to do is short)
__device__ __forceinline__ void key_body(const KeyArgs& a, uint32_t bid, uint32_t nb) {
    if (a.nlw && nl_reuse_prev(a.nlw)) return;  // the stage reuses its neighbour lists
    for (uint32_t base = bid * 256u; base < a.n; base += nb * 256u) {
        const uint32_t i = base + threadIdx.x;
        const bool in = i < a.n;
        uint32_t bin = 0, key = kInvalid;
        if (in) {
            float4 p = reinterpret_cast<const float4*>(a.posd)[i];
            if (a.cpyDstP) {  // the state kept only in the sorted buffers goes back to the landing ones
                reinterpret_cast<float4*>(a.cpyDstP)[i] = p;
                reinterpret_cast<float4*>(a.cpyDstV)[i] = reinterpret_cast<const float4*>(a.cpyVelq)[i];
            }
            uint32_t cx, cy, cz;
            if (cell_of(a.grid, p.x, p.y, p.z, cx, cy, cz)) {
                uint32_t row = cy + a.grid.sub[1] * cz;
                key = cx + a.grid.sub[0] * row;
                bin = row;
            } else {
                bin = a.grid.nrows;
            }
        }
        uint32_t rank = agg_rank(a.binCount, bin, in);
        if (in) {
            a.keys[i] = key;
            a.ranks[i] = rank;
        }
    }
}

// Exclusive scan of the row counts and of the occupied-row flags, one 1024-bin tile per
// workgroup with decoupled look-back; lists the occupied rows in ascending order and clears
// the counters for the next build.  A workgroup only waits on tiles whose ticket was taken
// before its own, and every tile publishes its aggregate before waiting, so each wait ends.
__device__ __forceinline__ unsigned long long ld_state(unsigned long long* p) {
    return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_state(unsigned long long* p, unsigned long long v) {
    __hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}
constexpr unsigned long long kStAgg = 1ull << 62, kStIncl = 2ull << 62, kStMask = 3ull << 62;

__device__ __forceinline__ void scan_body(const ScanArgs& a, uint32_t bid, uint32_t) {
    __shared__ uint32_t sTile, sEpoch;
    __shared__ uint32_t wsumC[4], wsumF[4];
    __shared__ unsigned long long sPrefix;
    const uint32_t tid = threadIdx.x;
    if (a.clear && bid == 0 && tid < a.nclear) a.clear[tid] = 0u;
    if (a.nlw) {
        const bool reuse = nl_reuse_prev(a.nlw);  // (read before the counter below advances)
        __syncthreads();
        if (bid == 0 && tid == 0) a.nlw[10] += 1u;  // the stage counter of the lists
        if (reuse) {  // the stage reuses its lists: no counts to scan
            if (bid == 0 && tid == 0) {
                if (a.stats) a.stats[0] = 0;
                if (a.stamp) *a.stamp += 1u;
            }
            return;
        }
    }
    if (tid == 0) {
        const uint32_t ep = __hip_atomic_load(a.epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        sEpoch = ep;
        sTile = atomicAdd(&a.ticket[ep & 1u], 1u);
    }
    __syncthreads();
    const uint32_t tile = sTile, par = sEpoch & 1u;
    unsigned long long* st = a.tileState + (size_t)par * a.ntiles;
    unsigned long long* stNext = a.tileState + (size_t)(par ^ 1u) * a.ntiles;
    const uint32_t nb = a.nbins;
    const uint32_t b0 = tile * kScanTile + tid * 4;
    uint32_t c[4], f[4], sc = 0, sf = 0;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const uint32_t b = b0 + k;
        c[k] = b < nb ? a.binCount[b] : 0u;
        f[k] = (c[k] > 0 && b + 1 < nb) ? 1u : 0u;
        sc += c[k];
        sf += f[k];
    }
    // block scan of (sc, sf)
    const uint32_t lane = __lane_id(), wid = tid >> 6;
    const uint32_t ic = wave_incl_scan(sc), iff = wave_incl_scan(sf);
    if (lane == 63) {
        wsumC[wid] = ic;
        wsumF[wid] = iff;
    }
    __syncthreads();
    uint32_t preC = 0, preF = 0;
    for (uint32_t w = 0; w < wid; w++) {
        preC += wsumC[w];
        preF += wsumF[w];
    }
    const uint32_t aggC = wsumC[0] + wsumC[1] + wsumC[2] + wsumC[3];
    const uint32_t aggF = wsumF[0] + wsumF[1] + wsumF[2] + wsumF[3];
    if (tid == 0) {
        const unsigned long long agg = ((unsigned long long)aggF << 32) | aggC;
        unsigned long long prefix = 0;
        if (tile == 0) {
            st_state(&st[0], kStIncl | agg);
        } else {
            st_state(&st[tile], kStAgg | agg);
            for (int64_t j = (int64_t)tile - 1; j >= 0;) {
                const unsigned long long v = ld_state(&st[j]);
                const unsigned long long s2 = v & kStMask;
                if (s2 == 0) continue;  // predecessor has not published yet
                prefix += v & ~kStMask;
                if (s2 == kStIncl) break;
                j--;
            }
            st_state(&st[tile], kStIncl | (prefix + agg));
        }
        sPrefix = prefix;
        stNext[tile] = 0;
        if (tile == 0) {
            a.ticket[par ^ 1u] = 0;
            if (a.stats) a.stats[0] = 0;
            if (a.stamp) *a.stamp += 1u;
        }
        if (tile == a.ntiles - 1) {
            const unsigned long long tot = prefix + agg;
            a.binStart[nb] = (uint32_t)(tot & 0xFFFFFFFFull);
            *a.numActive = (uint32_t)(tot >> 32);
            __hip_atomic_store(a.epoch, sEpoch + 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
    __syncthreads();
    uint32_t es = (uint32_t)(sPrefix & 0xFFFFFFFFull) + preC + ic - sc;
    uint32_t ef = (uint32_t)(sPrefix >> 32) + preF + iff - sf;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const uint32_t b = b0 + k;
        if (b < nb) {
            a.binStart[b] = es;
            if (f[k]) a.activeRows[ef++] = b;
            es += c[k];
            a.binCount[b] = 0;
        }
    }
}

__device__ __forceinline__ void move_particle(const ScatterArgs& a, uint32_t src, uint32_t dst) {
    reinterpret_cast<float4*>(a.dstPosd)[dst] = reinterpret_cast<const float4*>(a.srcPosd)[src];
    reinterpret_cast<float4*>(a.dstVelq)[dst] = reinterpret_cast<const float4*>(a.srcVelq)[src];
    a.dstMk[dst] = a.srcMk[src];
    uint32_t p = a.srcPid[src];
    a.dstPid[dst] = p;
    if (a.inv) a.inv[p] = dst;
    if (a.recMk) {
        a.recMk[dst] = a.srcMk[src];
        a.recPid[dst] = p;
    }
    if (a.dstAuxP) {
        reinterpret_cast<float4*>(a.dstAuxP)[dst] = reinterpret_cast<const float4*>(a.srcAuxP)[src];
        reinterpret_cast<float4*>(a.dstAuxV)[dst] = reinterpret_cast<const float4*>(a.srcAuxV)[src];
    }
}

// Whether two positions lie in the same or neighbouring buckets (27-bucket stencil,
// neighbor_list.rs:10-33 / fluid.rs:137-138).
__device__ __forceinline__ bool buckets_adjacent(const GridP& g, const float4& p, const float4& q) {
    uint32_t a0, a1, a2, b0, b1, b2;
    if (!cell_of(g, p.x, p.y, p.z, a0, a1, a2) || !cell_of(g, q.x, q.y, q.z, b0, b1, b2)) return false;
    return a0 + 1u >= b0 && b0 + 1u >= a0 && a1 + 1u >= b1 && b1 + 1u >= a1 && a2 + 1u >= b2 && b2 + 1u >= a2;
}

// Whether the stage about to run reuses the neighbour lists (see NlArgs).  Read after the
// stage's row-count scan advanced the counter; the words it reads are written only by the
// stage before.
// (Lists older than w[14] stages, if nonzero, are rebuilt: the build also re-sorts the
// storage, which the neighbour gathers' locality depends on.)
__device__ __forceinline__ bool nl_reuse(const uint32_t* w) {
    const uint32_t g = w[10], p = (g - 1u) & 1u;
    return w[p] == g - 1u && w[2 + p] != g - 1u && (w[14] == 0u || w[12 + p] + 1u < w[14]);
}
// The same decision before the stage's counter advances (key pass, scan): from the counter
// the previous stage recorded.
__device__ __forceinline__ bool nl_reuse_prev(const uint32_t* w) {
    const uint32_t g1 = w[6], p = g1 & 1u;
    return g1 != kInvalid && w[p] == g1 && w[2 + p] != g1 && (w[14] == 0u || w[12 + p] + 1u < w[14]);
}

// Scatter (key, source) into the reserved row slots; particles with non-finite positions
// go straight to their final place after all rows (neighbor_list.rs:86-93 drops them).
// A stage that reuses its neighbour lists keeps the storage order and places nothing (with
// copyOnReuse a full build copies the state to the sorted buffers).  Grid-stride over the
// particles, so that a launch with nothing to do is short.
__device__ __forceinline__ void scatter_one(const ScatterArgs& a, uint32_t i);
__device__ __forceinline__ void scatter_body(const ScatterArgs& a, uint32_t bid, uint32_t nb) {
    const bool reuse = a.nlw && nl_reuse(a.nlw);
    if (reuse && !(a.full && (a.copyOnReuse || a.recMk))) return;
    for (uint32_t i = bid * 256u + threadIdx.x; i < a.n; i += nb * 256u) {
        if (reuse) {
            if (a.copyOnReuse) {
                reinterpret_cast<float4*>(a.dstPosd)[i] = reinterpret_cast<const float4*>(a.srcPosd)[i];
                reinterpret_cast<float4*>(a.dstVelq)[i] = reinterpret_cast<const float4*>(a.srcVelq)[i];
                a.dstMk[i] = a.srcMk[i];
                a.dstPid[i] = a.srcPid[i];
                if (a.dstAuxP) {
                    reinterpret_cast<float4*>(a.dstAuxP)[i] = reinterpret_cast<const float4*>(a.srcAuxP)[i];
                    reinterpret_cast<float4*>(a.dstAuxV)[i] = reinterpret_cast<const float4*>(a.srcAuxV)[i];
                }
            }
            if (a.srcOut) a.srcOut[i] = i;
            if (a.recMk) {
                a.recMk[i] = a.srcMk[i];
                a.recPid[i] = a.srcPid[i];
            }
        } else {
            scatter_one(a, i);
        }
    }
}
__device__ __forceinline__ void scatter_one(const ScatterArgs& a, uint32_t i) {
    const uint32_t key = a.keys[i], rank = a.ranks[i];
    if (key != kInvalid) {
        uint32_t row = key / a.grid.sub[0];
        uint32_t slot = a.binStart[row] + rank;
        a.tmpKey[slot] = key;
        a.tmpSrc[slot] = i;
    } else {
        uint32_t dst = a.binStart[a.grid.nrows] + rank;
        if (a.full) {
            a.perm[dst] = dst | kFlag;
            move_particle(a, i, dst);
            if (a.srcOut) a.srcOut[dst] = i;
        } else {
            a.perm[dst] = i | kFlag;
        }
        if (a.sflag) a.sflag[a.full ? dst : i] = 1u;
    }
}

// One workgroup per occupied (y,z) row: LDS counting sort by x, then ascending particle id
// inside each cell (the order the 128-slot cap keeps), per-row cell offsets, cap flags,
// and the permutation / reorder of the state into cell-major storage.
__device__ __forceinline__ void rowsort_body(const RowSortArgs& a, uint32_t bid, uint32_t nb) {
    __shared__ uint32_t cnt[kMaxSX + 1];
    __shared__ uint32_t off[kMaxSX + 1];
    __shared__ uint32_t ePid[kRowCap];
    __shared__ uint32_t eGrp[kRowCap];
    __shared__ uint16_t eX[kRowCap];
    __shared__ uint32_t wsum[4];
    const uint32_t tid = threadIdx.x;
    if (a.nlw && nl_reuse(a.nlw)) return;  // the stage reuses its neighbour lists
    const uint32_t sx = a.grid.sub[0];
    const uint32_t nact = *a.numActive;
    for (uint32_t ri = bid; ri < nact; ri += nb) {
        const uint32_t row = a.activeRows[ri];
        const uint32_t s0 = a.binStart[row];
        const uint32_t ne = a.binStart[row + 1] - s0;
        const bool big = ne > (uint32_t)kRowCap;
        uint32_t* pidA = big ? a.gPid + s0 : ePid;
        uint32_t* grpA = big ? a.gGrp + s0 : eGrp;
        for (uint32_t x = tid; x <= sx; x += 256) cnt[x] = 0;
        __syncthreads();
        for (uint32_t e = tid; e < ne; e += 256) {
            uint32_t key = a.tmpKey[s0 + e];
            uint32_t x = key - row * sx;
            uint32_t src = a.tmpSrc[s0 + e];
            uint32_t p = a.srcPid[src];
            if (!big) eX[e] = (uint16_t)x;
            pidA[e] = p;
            atomicAdd(&cnt[x], 1u);
        }
        __syncthreads();
        // exclusive scan of cnt[0..sx) -> off; 8 consecutive cells per thread
        uint32_t loc[8];
        uint32_t sum = 0;
#pragma unroll
        for (int k = 0; k < 8; k++) {
            uint32_t idx = tid * 8 + k;
            uint32_t v = idx < sx ? cnt[idx] : 0;
            loc[k] = sum;
            sum += v;
        }
        uint32_t total;
        uint32_t ex = block_excl_scan256(sum, wsum, total);
#pragma unroll
        for (int k = 0; k < 8; k++) {
            uint32_t idx = tid * 8 + k;
            if (idx < sx) off[idx] = ex + loc[k];
        }
        if (tid == 0) off[sx] = total;
        __syncthreads();
        uint32_t* ro = a.rowOff + (size_t)row * a.grid.rowstride;
        for (uint32_t x = tid; x <= sx; x += 256) {
            ro[x] = s0 + off[x];
            if (x < sx) cnt[x] = 0;  // reused as fill counters
        }
        if (tid == 0) a.rowStamp[row] = *a.stamp;
        __syncthreads();
        for (uint32_t e = tid; e < ne; e += 256) {
            uint32_t x = big ? a.tmpKey[s0 + e] - row * sx : eX[e];
            uint32_t q = off[x] + atomicAdd(&cnt[x], 1u);
            grpA[q] = e;
        }
        __syncthreads();
        for (uint32_t q = tid; q < ne; q += 256) {
            const uint32_t e = grpA[q];
            const uint32_t x = big ? a.tmpKey[s0 + e] - row * sx : eX[e];
            const uint32_t p = pidA[e];
            const uint32_t c0 = off[x], c1 = off[x + 1];
            uint32_t r;
            if (c1 - c0 <= 256) {
                r = 0;
                for (uint32_t k = c0; k < c1; k++) r += pidA[grpA[k]] < p ? 1u : 0u;
            } else {
                r = q - c0;  // degenerate cell: keep arrival order
            }
            const uint32_t dst = s0 + c0 + r;
            const size_t cell = (size_t)row * sx + x;
            // slots left for mobile particles after the boundary and reference segments
            // (neighbor_list.rs:98-110): 128 - count[0] - count[1]
            const uint32_t used = (a.bP[cell + 1] - a.bP[cell]) + (a.rP[cell + 1] - a.rP[cell]);
            const uint32_t capn = kBucketCap - used;
            const uint32_t flag = r >= capn ? kFlag : 0u;
            const uint32_t src = a.tmpSrc[s0 + e];
            if (a.full) {
                a.perm[dst] = dst | flag;
                reinterpret_cast<float4*>(a.dstPosd)[dst] = reinterpret_cast<const float4*>(a.srcPosd)[src];
                reinterpret_cast<float4*>(a.dstVelq)[dst] = reinterpret_cast<const float4*>(a.srcVelq)[src];
                a.dstMk[dst] = a.srcMk[src];
                a.dstPid[dst] = p;
                if (a.inv) a.inv[p] = dst;
                if (a.recMk) {
                    a.recMk[dst] = a.srcMk[src];
                    a.recPid[dst] = p;
                }
                if (a.dstAuxP) {
                    reinterpret_cast<float4*>(a.dstAuxP)[dst] = reinterpret_cast<const float4*>(a.srcAuxP)[src];
                    reinterpret_cast<float4*>(a.dstAuxV)[dst] = reinterpret_cast<const float4*>(a.srcAuxV)[src];
                }
                if (a.srcOut) a.srcOut[dst] = src;
            } else {
                a.perm[dst] = src | flag;
            }
            if (a.sflag) a.sflag[a.full ? dst : src] = flag ? 1u : 0u;
            if (flag && a.stats) atomicAdd(a.stats, 1u);
        }
        __syncthreads();
    }
}

// ------------------------------------------------------------------------------------
// fluid_forces (fluid.rs:95-251) with the RK stage epilogue fused in
// ------------------------------------------------------------------------------------
constexpr int kList = 24;  // within-h neighbours buffered per thread before a drain

struct PairAcc {
    float den, fx, fy, fz;
};

// Stage epilogue of one particle: y_n + dt*sum a k -> next stage input (or y_{n+1}), the
// b-weighted accumulator, stored derivatives, the state's p/den^2 and the next cell key.
__device__ __forceinline__ void rk_update_particle(const RkArgs& r, const GridP& g, const DMat& M,
                                                   uint32_t i, float4 kP, float4 kV, bool& haveKey,
                                                   uint32_t& nextKey, uint32_t& nextBin, float4& opOut) {
    const float4 yp = reinterpret_cast<const float4*>(r.Yposd)[i];
    const float4 yv = reinterpret_cast<const float4*>(r.Yvelq)[i];
    float4 ap = make_float4(0.f, 0.f, 0.f, 0.f), av = ap;
    if (r.AccP) {
        ap = reinterpret_cast<const float4*>(r.AccP)[i];
        av = reinterpret_cast<const float4*>(r.AccV)[i];
    }
    const float b = r.b_cur, bp = r.b_pos;
    ap.x += bp * kP.x; ap.y += bp * kP.y; ap.z += bp * kP.z; ap.w += b * kP.w;
    av.x += b * kV.x; av.y += b * kV.y; av.z += b * kV.z;
    float4 op, ov;
    if (r.last) {
        op = make_float4(yp.x + ap.x, yp.y + ap.y, yp.z + ap.z, yp.w + ap.w);
        ov = make_float4(yv.x + av.x, yv.y + av.y, yv.z + av.z, 0.f);
    } else {
        const float c = r.a_next;
        op = make_float4(yp.x + c * kP.x, yp.y + c * kP.y, yp.z + c * kP.z, yp.w + c * kP.w);
        ov = make_float4(yv.x + c * kV.x, yv.y + c * kV.y, yv.z + c * kV.z, 0.f);
        for (int kk = 0; kk < r.nk; kk++) {
            const float4 p2 = reinterpret_cast<const float4*>(r.Kp[kk])[i];
            const float4 v2 = reinterpret_cast<const float4*>(r.Kv[kk])[i];
            const float ck = r.kc[kk];
            op.x += ck * p2.x; op.y += ck * p2.y; op.z += ck * p2.z; op.w += ck * p2.w;
            ov.x += ck * v2.x; ov.y += ck * v2.y; ov.z += ck * v2.z;
        }
        if (r.AccPOut) {
            reinterpret_cast<float4*>(r.AccPOut)[i] = ap;
            reinterpret_cast<float4*>(r.AccVOut)[i] = av;
        }
    }
    if (r.KpOut) {
        reinterpret_cast<float4*>(r.KpOut)[i] = kP;
        reinterpret_cast<float4*>(r.KvOut)[i] = make_float4(kV.x, kV.y, kV.z, 0.0f);
    }
    ov.w = pressure_q(M, op.w);
    opOut = op;
    if (r.vdt != 0.0f) {  // velocity Verlet: the next force point (as k_verlet_pre forms it)
        const float dt = r.vdt, h = 0.5f * dt;
        opOut.x = op.x + dt * (ov.x + h * kV.x);
        opOut.y = op.y + dt * (ov.y + h * kV.y);
        opOut.z = op.z + dt * (ov.z + h * kV.z);
    }
    reinterpret_cast<float4*>(r.Oposd)[i] = op;
    if (r.Oposd2) {
        reinterpret_cast<float4*>(r.Oposd2)[i] = op;
        reinterpret_cast<float4*>(r.Ovelq2)[i] = ov;
    }
    reinterpret_cast<float4*>(r.Ovelq)[i] = ov;
    if (r.keys) {
        uint32_t cx, cy, cz;
        haveKey = true;
        if (cell_of(g, op.x, op.y, op.z, cx, cy, cz)) {
            const uint32_t row = cy + g.sub[1] * cz;
            nextKey = cx + g.sub[0] * row;
            nextBin = row;
        } else {
            nextKey = kInvalid;
            nextBin = g.nrows;
        }
    }
}

// The same epilogue on a solid's plastic strain (16 floats, solid-id order).
__device__ __forceinline__ void rk_update_eps(const RkArgs& r, uint32_t s, const float* ke) {
    const float* y = r.Yeps + 16 * (size_t)s;
    float out[16], acc[16];
    for (int q = 0; q < 16; q++) acc[q] = r.AccE ? r.AccE[16 * (size_t)s + q] : 0.0f;
    for (int q = 0; q < 16; q++) acc[q] += r.b_cur * ke[q];
    if (r.last) {
        for (int q = 0; q < 16; q++) out[q] = y[q] + acc[q];
    } else {
        for (int q = 0; q < 16; q++) out[q] = y[q] + r.a_next * ke[q];
        for (int kk = 0; kk < r.nk; kk++)
            for (int q = 0; q < 16; q++) out[q] += r.kc[kk] * r.Ke[kk][16 * (size_t)s + q];
        if (r.AccEOut)
            for (int q = 0; q < 16; q++) r.AccEOut[16 * (size_t)s + q] = acc[q];
    }
    if (r.KeOut)
        for (int q = 0; q < 16; q++) r.KeOut[16 * (size_t)s + q] = ke[q];
    for (int q = 0; q < 16; q++) r.Oeps[16 * (size_t)s + q] = out[q];
}

// Logical block id that gives each XCD a contiguous share of the grid: workgroups are
// dispatched round-robin over the 8 XCDs, each with its own L2, and neighbouring particles
// then share an L2.
__device__ __forceinline__ uint32_t xcd_block() {
    const uint32_t nb = gridDim.x, b = blockIdx.x;
    const uint32_t q = nb >> 3, r = nb & 7u, x = b & 7u, k = b >> 3;
    return (x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q) + k;
}

struct PairIn {
    float4 xj, vj;
    uint32_t mk2;
    bool bnd;
};

#ifndef ARP_ADJ_GROUP
#define ARP_ADJ_GROUP 4  // solid adjoint list entries whose gathers are issued together
#endif
// per-solid F and F' cotangents of the solid reverse pass: 18 floats, padded to 20 (float4 rows)
constexpr uint32_t kFbarStride = 20;
#ifndef ARP_ST_GROUP
#define ARP_ST_GROUP 4  // strain list entries whose gathers are issued together
#endif
#ifndef ARP_EL_GROUP
#define ARP_EL_GROUP 4  // elastic-force list entries whose gathers are issued together
#endif

#ifndef ARP_BATCH
#define ARP_BATCH 4  // candidates whose loads are issued together (arrays are padded by this)
#endif

// LDS of the fluid kernels (namespace scope, so that both table variants of a kernel share it)
__shared__ uint32_t s_list[kList * 256];     // candidates buffered per thread
__shared__ float4 s_mat[kMatLds];            // mass, visc, bulk_visc
__shared__ DInter s_inter[kMatLds * kMatLds];

#ifndef ARP_NLB
#define ARP_NLB 4  // list entries whose neighbour loads are issued together (list walk)
#endif
#ifndef ARP_WALK_DIRECT
// list walk without LDS compaction: the pair follows the test in registers (fewer gather
// requests: no reload of the position, no velocity request for entries beyond h)
#define ARP_WALK_DIRECT 1
#endif
#ifndef ARP_PP_ALL
#define ARP_PP_ALL 0  // two neighbour register sets in the search kernels' drains too
#endif

// MODE 0: plain step.  MODE 1: also keep every particle's neighbours within h (the recording
// of the reverse sweep).  MODES 2, 3: neighbour lists kept across stages (NlArgs), one
// kernel each, both launched for every stage; the one that does not apply to the stage
// returns at once.  MODE 2 searches the grid with radius h + skin and writes new lists
// (the entries found within h first, the others from the end of the particle's list);
// MODE 3 walks the kept lists.  MODES 4, 5: 2, 3 keeping, as MODE 1 does, each particle's
// neighbours within h.  Separate instances, so each carries only its own registers.
// LT: the material and interaction tables are read from LDS (at most kMatLds materials);
// a compile-time choice, so that the pair loop's table reads are LDS reads, never generic
// (flat) loads, whose completion the compiler could only await together with the
// neighbour loads already in flight.
// FM (walks only): every mobile particle has material fastMat, every boundary particle
// fastBnd, and none is elastic; a pair's material parameters are then selected by whether
// the neighbour is a boundary particle, without per-neighbour material loads.
template <int MODE, bool LT, bool FM = false, bool DIRECT = (ARP_WALK_DIRECT != 0)>
__device__ __forceinline__ void fluid_body(const FluidArgs& a, uint32_t bid, uint32_t) {
    static_assert(!FM || MODE == 3 || MODE == 5, "the fast-material form walks kept lists");
    constexpr bool REC = MODE == 1 || MODE >= 4;
    constexpr bool NL = MODE >= 2;
    constexpr bool WALK = MODE == 3 || MODE == 5;
    const uint32_t tid = threadIdx.x;
    // MODES 2, 3: the stage's build counter and whether it walks the kept lists
    uint32_t gen = 0;
    bool use = false;
    bool swalk = false, refresh = false;  // two-level lists: short walk, short-list refresh
    const bool spec = WALK && a.spec != 0;
    if (NL) {
        if (spec) {
            // a speculative stage: its counter value comes from the host, and the decision
            // reads only the previous stage's words (the other parity from those this kernel
            // writes), so every block takes the same one
            if (a.nl.words[kNlAbort] != kInvalid) return;  // an earlier stage was not valid
            gen = a.specGen;
            {
                const uint32_t* w = a.nl.words;
                const uint32_t p = (gen - 1u) & 1u;
                use = w[p] == gen - 1u && w[2 + p] != gen - 1u && (w[14] == 0u || w[12 + p] + 1u < w[14]);
            }
            if (!use) {
                if (bid == 0 && tid == 0) {
                    a.nl.words[kNlAbort] = a.specStep;
                    __hip_atomic_store(a.hostAbort, a.specStep, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
                    __threadfence_system();
                }
                return;
            }
            if (bid == 0 && tid == 0) {
                if (a.specSnap) {  // the words as the step found them
                    for (int q = 0; q < 15; q++) a.specSnap[q] = a.nl.words[q];
                    a.specSnap[15] = *a.stampW;
                    a.specSnap[16] = a.statsW[0];
                    for (int q = 16; q < 19; q++) a.specSnap[q + 1] = a.nl.words[q];
                }
                // the bookkeeping of the rebuild chain's reusing stage (no block reads these)
                a.nl.words[10] = gen;
                a.statsW[0] = 0u;
                *a.stampW += 1u;
                // a recorded step's flags start at zero (a walk never overflows the record,
                // whose capacity is at least the lists', so no other block writes them)
                if (REC && a.specClear)
                    for (int q = 0; q < kRecFlags; q++) a.specClear[q] = 0u;
            }
            if (REC && a.recSrc) {  // the reusing sort's records: storage order unchanged
                const uint32_t t0 = bid * 256u + tid;
                if (t0 < a.n) {
                    a.recSrc[t0] = t0;
                    a.recMk[t0] = a.mk[t0];
                    a.recPid[t0] = a.pidIn[t0];
                }
            }
        } else {
            gen = a.nl.words[10];
            use = nl_reuse(a.nl.words);
            if (use != WALK) return;  // the other kernel of the pair runs this stage
        }
        // two-level lists: a walking stage uses the short list unless the previous stage found
        // a particle beyond its reach; then it walks the long list and writes a new short one
        if (use && a.nl.srad2 > 0.0f) {
            swalk = a.nl.words[16 + ((gen - 1u) & 1u)] != gen - 1u;
            refresh = !swalk;
        }
        if (bid == 0 && tid == 0) {
            if (REC) *a.recUse = use ? 1u : 0u;
            if (refresh) a.nl.words[18] += 1u;
            a.nl.words[gen & 1u] = gen;
            a.nl.words[6] = gen;
            a.nl.words[12 + (gen & 1u)] = use ? a.nl.words[12 + ((gen - 1u) & 1u)] + 1u : 0u;  // age
            atomicAdd(&a.nl.words[use ? 9 : 8], 1u);
            if (!use) {
                a.nl.words[4 + (gen & 1u)] = gen;
                // a build over a sort that refused particles (bucket cap) is not kept
                if (a.stats && *a.stats != 0u) a.nl.words[2 + (gen & 1u)] = gen;
            }
        }
    }
    uint32_t* list = s_list;
    float4* smat = s_mat;
    DInter* sint = s_inter;
    const uint32_t nm = a.phys.nm;
    if (LT) {
        for (uint32_t m = tid; m < nm; m += 256)
            smat[m] = make_float4(a.mats[m].mass, a.mats[m].visc, a.mats[m].bulk_visc, 0.f);
        for (uint32_t m = tid; m < nm * nm; m += 256) sint[m] = a.inter[m];
        __syncthreads();
    }
    const uint32_t t = bid * 256u + tid;
    const bool act = t < a.n;
    const float4* X = reinterpret_cast<const float4*>(a.Xposd);
    const float4* XV = reinterpret_cast<const float4*>(a.Xvelq);
    const float4* B = reinterpret_cast<const float4*>(a.bposd);
    const float4* BV = reinterpret_cast<const float4*>(a.bvelq);
    const float h = a.phys.h, hh = h * h, kn = a.phys.k;
    // candidates are tested against this radius (MODE 2 builds: h + skin)
    const float testR2 = (NL && !use) ? a.nl.rad2 : hh;

    uint32_t i = 0, nextBin = 0, nextKey = kInvalid;
    bool haveKey = false;
    if (act) {
        bool noNb;
        if (NL && use) {
            i = t;  // the lists index storage order, which a reused stage keeps
            noNb = false;
        } else if (a.identity) {
            i = t;
            noNb = a.sflag[t] != 0u;
        } else {
            const uint32_t pe = a.perm[t];
            i = pe & ~kFlag;
            noNb = (pe & kFlag) != 0;
        }
        const float4 xi = X[i];
        const float4 vi = XV[i];
        const uint32_t mk1 = FM ? a.fastMat : a.mk[i];
        const uint32_t mat1 = mk1 & kMatMask;
        const bool elastic1 = FM ? false : (mk1 & kMkElastic) != 0;
        const DMat& M1 = a.mats[mat1];
        const float m1 = M1.mass, d1 = xi.w, q1 = vi.w;
        const float f1 = M1.visc, bf1 = M1.bulk_visc;
        PairAcc acc = {0.f, 0.f, 0.f, 0.f};
        // FM: the two neighbour kinds' parameters (uniform, loaded once)
        float fmM[2] = {0.f, 0.f}, fmV[2] = {0.f, 0.f}, fmB[2] = {0.f, 0.f};
        DInter fmI[2] = {};
        if constexpr (FM) {
            const uint32_t mm[2] = {a.fastMat, a.fastBnd};
#pragma unroll
            for (int k = 0; k < 2; k++) {
                fmM[k] = a.mats[mm[k]].mass;
                fmV[k] = a.mats[mm[k]].visc;
                fmB[k] = a.mats[mm[k]].bulk_visc;
                fmI[k] = a.inter[mat1 * nm + mm[k]];
            }
        }

        auto load = [&](uint32_t e) {
            PairIn in;
            in.bnd = (e & kBnd) != 0;
            const uint32_t j = e & ~kBnd;
            in.xj = (in.bnd ? B : X)[j];
            in.vj = (in.bnd ? BV : XV)[j];
            in.mk2 = FM ? (in.bnd ? a.fastBnd : a.fastMat) : (in.bnd ? a.bmk : a.mk)[j];
            return in;
        };
        // one neighbour's terms (fluid.rs:153-214)
        auto pair = [&](const PairIn& in) {
            const bool bnd = in.bnd;
            const float4 xj = in.xj, vj = in.vj;
            const uint32_t mat2 = in.mk2 & kMatMask;
            const bool elastic2 = FM ? false : (in.mk2 & kMkElastic) != 0;
            float m2, visc2, bulk2;
            DInter it;
            if constexpr (FM) {
                const int k = bnd ? 1 : 0;
                m2 = bnd ? fmM[1] : fmM[0];
                visc2 = bnd ? fmV[1] : fmV[0];
                bulk2 = bnd ? fmB[1] : fmB[0];
                it.potential = bnd ? fmI[1].potential : fmI[0].potential;
                it.strength = bnd ? fmI[1].strength : fmI[0].strength;
                it.radius = bnd ? fmI[1].radius : fmI[0].radius;
                it.dampening = bnd ? fmI[1].dampening : fmI[0].dampening;
                it.friction = bnd ? fmI[1].friction : fmI[0].friction;
                (void)k;
            } else if constexpr (LT) {
                const float4 mm = smat[mat2];
                m2 = mm.x;
                visc2 = mm.y;
                bulk2 = mm.z;
                it = sint[mat1 * nm + mat2];
            } else {
                m2 = a.mats[mat2].mass;
                visc2 = a.mats[mat2].visc;
                bulk2 = a.mats[mat2].bulk_visc;
                it = a.inter[mat1 * nm + mat2];
            }
            const float d2 = xj.w, q2 = vj.w;
            const float rx = xj.x - xi.x, ry = xj.y - xi.y, rz = xj.z - xi.z;
            const float vx = vj.x - vi.x, vy = vj.y - vi.y, vz = vj.z - vi.z;
            const float rr = rx * rx + ry * ry + rz * rz;
            // a build's list also holds candidates between h and h + skin: no terms
            if (NL && !WALK && !(hh - rr > 0.0f)) return;
            // the reference pairs only particles in neighbouring buckets; a pair within h can
            // be two buckets apart only within the rounding of the bucket function of h
            if (NL && rr > a.nl.grey2) {
                // (rare: the buckets are recomputed here, not kept across the loop)
                float4 pi = xi;
                asm volatile("" : "+v"(pi.x), "+v"(pi.y), "+v"(pi.z));
                if (!buckets_adjacent(a.grid, pi, xj)) return;
            }
            const float diff = fmaxf(hh - rr, 0.0f);
            const float gs = -6.0f * kn * diff * diff;
            const float gx = gs * rx, gy = gs * ry, gz = gs * rz;
            const float vr = vx * rx + vy * ry + vz * rz;
            if (!bnd) acc.den += m2 * (gs * vr);
            const float cp = it.potential ? contact_potential(it.potential, sqrtf(rr), it.radius, it.strength) : 0.0f;
            // pressure along grad W; a mobile pair's artificial viscosity lies along it too and
            // joins the same coefficient below
            float pf = m1 * m2 * (q1 + q2);
            if (cp != 0.0f) {
                const float cs = fdiv(m2 * cp, d2);
                const float cx = cs * gx, cy = cs * gy, cz = cs * gz;
                if ((elastic1 || elastic2 || bnd) && mat1 != mat2 && it.friction != 0.0f) {
                    // friction along the tangential relative velocity (fluid.rs:191-199)
                    const float irr = fdiv(1.0f, rr);
                    const float ix = rx * irr, iy = ry * irr, iz = rz * irr;
                    const float cfr = cx * rx + cy * ry + cz * rz;
                    const float nx = ix * cfr, ny = iy * cfr, nz = iz * cfr;
                    float tx = vx - ix * vr, ty = vy - iy * vr, tz = vz - iz * vr;
                    const float itl = fdiv(1.0f, sqrtf(tx * tx + ty * ty + tz * tz));
                    tx *= itl;
                    ty *= itl;
                    tz *= itl;
                    if (isfinite(tx) && isfinite(ty) && isfinite(tz)) {
                        const float fl = it.friction * sqrtf(nx * nx + ny * ny + nz * nz);
                        const float itl2 = fdiv(fl, sqrtf(tx * tx + ty * ty + tz * tz));
                        acc.fx += itl2 * tx;
                        acc.fy += itl2 * ty;
                        acc.fz += itl2 * tz;
                    }
                }
                pf += cs;  // (the contact force cs grad W)
            }
            if (!bnd) {
                const float den_ = rr + a.phys.eps_hh;
                if (!elastic1 && f1 + visc2 != 0.0f) {
                    const float s = fdiv(m1 * m2 * (f1 + visc2) * (rx * gx + ry * gy + rz * gz), d1 * d2 * den_);
                    acc.fx -= s * vx;
                    acc.fy -= s * vy;
                    acc.fz -= s * vz;
                }
                const float ff = mat1 == mat2 ? bf1 + bulk2 : it.dampening;
                pf -= m1 * m2 * fdiv(ff * vr, (d1 + d2) * den_);
            }
            acc.fx += pf * gx;
            acc.fy += pf * gy;
            acc.fz += pf * gz;
        };
        // the buffered candidates' terms; the next neighbour's loads are issued before the
        // current pair's arithmetic, into the other of two register sets (no copies, so a
        // pair waits only for its own loads)
        auto drain = [&](uint32_t nl) {
            if (WALK || ARP_PP_ALL) {
                // (the next loads are unconditional: a lane at its last pair reloads its
                // first entry rather than branching around the load)
                PairIn pa, pb;
                if (nl) pa = load(list[tid]);
                for (uint32_t k = 0; k < nl; k += 2) {
                    pb = load(list[(k + 1 < nl ? k + 1 : 0u) * 256 + tid]);
                    pair(pa);
                    if (k + 1 >= nl) break;
                    pa = load(list[(k + 2 < nl ? k + 2 : 0u) * 256 + tid]);
                    pair(pb);
                }
            } else {  // one register set (the search kernels keep five waves per SIMD)
                PairIn nxt;
                if (nl) nxt = load(list[tid]);
                for (uint32_t k = 0; k < nl; k++) {
                    const PairIn cur2 = nxt;
                    if (k + 1 < nl) nxt = load(list[(k + 1) * 256 + tid]);
                    pair(cur2);
                }
            }
        };

        uint32_t nrec = 0;  // neighbours recorded (reverse sweep)
        uint32_t nfront = 0, nback = 0;  // list entries within h / within h + skin (MODE 2)
        // two-level lists: the short list written by a build or a refresh, entries within h
        // at its front, the others from its end (as the long list)
        uint32_t sfront = 0, sback = 0;
        auto short_push = [&](uint32_t v, bool inh) {
            const uint32_t cap = a.nl.cap;
            uint32_t& c = inh ? sfront : sback;
            if (c < cap) a.nl.slist[(size_t)(inh ? c : cap - 1u - c) * a.n + i] = v;
            c++;
        };
        if constexpr (WALK) {
            // walk the kept list, ARP_NLB entries per batch: the direct walk evaluates the
            // entries now within h as they are tested; the recording walk buffers them in LDS
            // (branch-free test), records them and drains them as the search does
            const uint32_t cw = (swalk ? a.nl.scount : a.nl.count)[i];
            const uint32_t fa = cw & 0xFFFFu, fb = cw >> 16, m = fa + fb, cap = a.nl.cap;
            const uint32_t* L = swalk ? a.nl.slist : a.nl.list;
            // 32-bit element offsets (the host keeps cap x n below 2^32): entry k at k n + i,
            // the back entries (k >= fa) cap - fb - fa further on
            const uint32_t n32 = a.n, jump = (cap - fb - fa) * n32;
            uint32_t k = 0, nl = 0;
            // the entries k0 .. k0 + ARP_NLB - 1 (unconditional loads: a lane past its count
            // reads its entry 0; a 32-bit byte offset from the uniform base, an SGPR-based load)
            auto fetch = [&](uint32_t k0, uint32_t (&eo)[ARP_NLB], uint32_t (&oo)[ARP_NLB]) {
#pragma unroll
                for (int u = 0; u < ARP_NLB; u++) {
                    const uint32_t kk = k0 + u;
                    oo[u] = kk < m ? 1u : 0u;
                    const uint32_t off = oo[u] ? i + kk * n32 + (kk >= fa ? jump : 0u) : i;
                    const uint32_t v = *reinterpret_cast<const uint32_t*>(reinterpret_cast<const char*>(L) + off * 4u);
                    eo[u] = oo[u] ? v : 0u;
                }
            };
            // the entries of the next batch are loaded while the current batch's neighbour
            // positions are in flight, so each batch waits for one round trip, not two
            uint32_t e[ARP_NLB], ok[ARP_NLB];
            fetch(0, e, ok);
            if constexpr (DIRECT) {
                // direct walk: each entry's position is tested in registers and, for the
                // entries within h, the velocity is loaded (lanes outside h issue no request)
                // and the pair evaluated at once, in list order (the buffered walk's order);
                // a recording walk records those entries in the same order
                char* nbw = REC ? reinterpret_cast<char*>(a.nbList) : nullptr;
                for (; k < m; k += ARP_NLB) {
                    float4 xj[ARP_NLB];
#pragma unroll
                    for (int u = 0; u < ARP_NLB; u++) xj[u] = ((e[u] & kBnd) ? B : X)[e[u] & ~kBnd];
                    uint32_t en[ARP_NLB], okn[ARP_NLB];
                    fetch(k + ARP_NLB, en, okn);
                    bool inh[ARP_NLB];
                    PairIn in[ARP_NLB];
#pragma unroll
                    for (int u = 0; u < ARP_NLB; u++) {
                        const float rx = xj[u].x - xi.x, ry = xj[u].y - xi.y, rz = xj[u].z - xi.z;
                        const float rr = rx * rx + ry * ry + rz * rz;
                        inh[u] = ok[u] && hh - rr > 0.0f;
                        if (refresh && ok[u] && a.nl.srad2 - rr > 0.0f) short_push(e[u], inh[u]);
                        in[u].bnd = (e[u] & kBnd) != 0;
                        in[u].xj = xj[u];
                        const uint32_t j = e[u] & ~kBnd;
                        if (inh[u]) {
                            in[u].vj = (in[u].bnd ? BV : XV)[j];
                            in[u].mk2 = FM ? (in[u].bnd ? a.fastBnd : a.fastMat) : (in[u].bnd ? a.bmk : a.mk)[j];
                        }
                    }
#pragma unroll
                    for (int u = 0; u < ARP_NLB; u++)
                        if (inh[u]) {
                            if (REC) {
                                // (32-bit byte offsets: the host keeps the record below 4 GB)
                                if (nrec < a.nbCap)
                                    *reinterpret_cast<uint32_t*>(nbw + (nrec * n32 + i) * 4u) = e[u];
                                nrec++;
                            }
                            pair(in[u]);
                        }
#pragma unroll
                    for (int u = 0; u < ARP_NLB; u++) {
                        e[u] = en[u];
                        ok[u] = okn[u];
                    }
                }
            } else
            for (;;) {
                while (nl + ARP_NLB <= (uint32_t)kList && k < m) {
                    float4 xj[ARP_NLB];
#pragma unroll
                    for (int u = 0; u < ARP_NLB; u++) xj[u] = ((e[u] & kBnd) ? B : X)[e[u] & ~kBnd];
                    uint32_t en[ARP_NLB], okn[ARP_NLB];
                    fetch(k + ARP_NLB, en, okn);
#pragma unroll
                    for (int u = 0; u < ARP_NLB; u++) {
                        const float rx = xj[u].x - xi.x, ry = xj[u].y - xi.y, rz = xj[u].z - xi.z;
                        const float rr = rx * rx + ry * ry + rz * rz;
                        const uint32_t in = ok[u] & (hh - rr > 0.0f ? 1u : 0u);
                        if (refresh && ok[u] && a.nl.srad2 - rr > 0.0f) short_push(e[u], in != 0u);
                        list[nl * 256 + tid] = e[u];
                        nl += in;
                    }
#pragma unroll
                    for (int u = 0; u < ARP_NLB; u++) {
                        e[u] = en[u];
                        ok[u] = okn[u];
                    }
                    k += ARP_NLB;
                }
                if (REC) {  // the buffered entries are the neighbours within h
                    // (32-bit byte offsets from the uniform base: the host keeps the record
                    // below 4 GB)
                    const uint32_t nw = nrec < a.nbCap ? min(nl, a.nbCap - nrec) : 0u;
                    char* nbw = reinterpret_cast<char*>(a.nbList);
                    uint32_t off = (nrec * n32 + i) * 4u;
                    for (uint32_t q = 0; q < nw; q++, off += n32 * 4u)
                        *reinterpret_cast<uint32_t*>(nbw + off) = list[q * 256 + tid];
                    nrec += nl;
                }
                drain(nl);
                nl = 0;
                if (k >= m) break;
            }
        } else if (!noNb) {
            uint32_t cx, cy, cz;
            cell_of(a.grid, xi.x, xi.y, xi.z, cx, cy, cz);
            const uint32_t sx = a.grid.sub[0], sy = a.grid.sub[1], sz = a.grid.sub[2];
            const uint32_t stampv = *a.stamp;
            // stencil: the 3x3x3 buckets around the particle's (cell size >= h); a MODE 2
            // build covers every bucket within h + skin (the cells of the box x +- rad)
            uint32_t xl, xh;
            int ylo = (int)cy - 1, zlo = (int)cz - 1, nyr = 3, nrw = 9;
            if (NL) {
                const GridP& gp = a.grid;
                const float R = a.nl.rad;
                xl = cell_coord(xi.x - R, gp.min[0], gp.dim[0], sx);
                xh = cell_coord(xi.x + R, gp.min[0], gp.dim[0], sx);
                ylo = (int)cell_coord(xi.y - R, gp.min[1], gp.dim[1], sy);
                zlo = (int)cell_coord(xi.z - R, gp.min[2], gp.dim[2], sz);
                nyr = (int)cell_coord(xi.y + R, gp.min[1], gp.dim[1], sy) - ylo + 1;
                nrw = nyr * ((int)cell_coord(xi.z + R, gp.min[2], gp.dim[2], sz) - zlo + 1);
            } else {
                xl = cx > 0 ? cx - 1 : 0;
                xh = min(cx + 1, sx - 1);
            }
            // metadata of one neighbour row: [mobile lo, hi), [boundary lo, hi); all four
            // loads are independent, and the next row's are issued before this row is scanned
            auto row_meta = [&](int rw, uint4& m) {
                m = make_uint4(0, 0, 0, 0);
                const int nz = NL ? zlo + rw / nyr : (int)cz + rw / 3 - 1;
                const int ny = NL ? ylo + rw % nyr : (int)cy + rw % 3 - 1;
                if (rw < nrw && nz >= 0 && nz < (int)sz && ny >= 0 && ny < (int)sy) {
                    const uint32_t row = (uint32_t)ny + sy * (uint32_t)nz;
                    const uint32_t* ro = a.rowOff + (size_t)row * a.grid.rowstride;
                    const size_t cb = (size_t)row * sx;
                    const uint32_t st = a.rowStamp[row];
                    const uint32_t lo = ro[xl], hi = ro[xh + 1];
                    m.z = a.bP[cb + xl];
                    m.w = a.bP[cb + xh + 1];
                    if (st == stampv) {
                        m.x = lo;
                        m.y = hi;
                    }
                }
            };
            // Candidates within the test radius are buffered in LDS and drained at a single
            // site of the pair loop whenever the buffer could overflow.  Inside a batch the
            // test and the append are branch-free: every candidate is written to slot nl,
            // which advances only for accepted ones.
            uint32_t nl = 0;
            int rw = 0;
            uint4 cur, nxtm;
            row_meta(0, cur);
            row_meta(1, nxtm);
            int part = 0;  // 0: mobile segment, 1: boundary segment
            uint32_t q = cur.x, qend = cur.y;
            // a build without refused particles leaves sflag all zero (non-finite particles
            // are in no row), so the flag loads of the identity path can be skipped
            const bool flags = a.identity && (a.stats == nullptr || *a.stats != 0u);
            for (;;) {
                while (nl + ARP_BATCH <= (uint32_t)kList) {
                    if (q >= qend) {
                        if (part == 0) {
                            part = 1;
                            q = cur.z;
                            qend = cur.w;
                            continue;
                        }
                        if (++rw >= nrw) break;
                        cur = nxtm;
                        row_meta(rw + 1, nxtm);
                        part = 0;
                        q = cur.x;
                        qend = cur.y;
                        continue;
                    }
                    const bool bseg = part != 0;
                    const float4* base = bseg ? B : X;
                    const uint32_t tag = bseg ? kBnd : 0u;
                    uint32_t e[ARP_BATCH], ok[ARP_BATCH];
#pragma unroll
                    for (int u = 0; u < ARP_BATCH; u++) {
                        const uint32_t qq = q + u;  // < qend + ARP_BATCH: inside the padding
                        const uint32_t inr = qq < qend ? 1u : 0u;
                        uint32_t v = qq, okv = inr;
                        // uniform branches; the loads are unconditional (index 0 on boundary
                        // segments, whose indices may exceed the mobile arrays)
                        if (!a.identity) {
                            const uint32_t pe = a.perm[bseg ? 0u : qq];  // past qend: padding
                            okv = inr & ((bseg || !(pe & kFlag)) ? 1u : 0u);
                            v = bseg ? qq : (inr ? (pe & ~kFlag) : 0u);
                        } else if (flags) {
                            const uint32_t fl = a.sflag[bseg ? 0u : qq];
                            okv = inr & ((bseg || fl == 0u) ? 1u : 0u);
                        }
                        e[u] = v;
                        ok[u] = okv;
                    }
                    float4 xj[ARP_BATCH];
#pragma unroll
                    for (int u = 0; u < ARP_BATCH; u++) xj[u] = base[e[u]];
#pragma unroll
                    for (int u = 0; u < ARP_BATCH; u++) {
                        const float rx = xj[u].x - xi.x, ry = xj[u].y - xi.y, rz = xj[u].z - xi.z;
                        const uint32_t v = e[u] | tag;
                        // the self pair contributes exactly zero (grad W(0) = 0)
                        const uint32_t in = ok[u] & (v != i ? 1u : 0u) &
                                            (testR2 - (rx * rx + ry * ry + rz * rz) > 0.0f ? 1u : 0u);
                        list[nl * 256 + tid] = v;
                        nl += in;
                    }
                    q += ARP_BATCH;
                }
                if (REC && !NL) {
                    for (uint32_t k = 0; k < nl; k++)
                        if (nrec + k < a.nbCap) a.nbList[(size_t)(nrec + k) * a.n + i] = list[k * 256 + tid];
                    nrec += nl;
                }
                if (NL) {
                    // the candidates within h go to the front of the particle's list, the
                    // others (h .. h + skin) to its end
                    const uint32_t cap = a.nl.cap;
                    for (uint32_t k = 0; k < nl; k++) {
                        const uint32_t v = list[k * 256 + tid];
                        const uint32_t j = v & ~kBnd;
                        const float4 xj = ((v & kBnd) ? B : X)[j];
                        const float rx = xj.x - xi.x, ry = xj.y - xi.y, rz = xj.z - xi.z;
                        const float rr = rx * rx + ry * ry + rz * rz;
                        if (a.nl.srad2 - rr > 0.0f) short_push(v, hh - rr > 0.0f);
                        if (hh - rr > 0.0f) {
                            if (nfront < cap) a.nl.list[(size_t)nfront * a.n + i] = v;
                            nfront++;
                            if (REC) {
                                if (nrec < a.nbCap) a.nbList[(size_t)nrec * a.n + i] = v;
                                nrec++;
                            }
                        } else {
                            if (nback < cap) a.nl.list[(size_t)(cap - 1u - nback) * a.n + i] = v;
                            nback++;
                        }
                    }
                }
                drain(nl);
                nl = 0;
                if (rw >= nrw) break;
            }
        }
        if (REC) {
            a.nbCount[i] = min(nrec, a.nbCap);
            if (nrec > a.nbCap) atomicOr(a.nbOver, 1u);
        }
        if (NL) {
            // clear the bucket counts of the previous stage's build (the buckets recorded at
            // this storage index; every index is visited once, so every count is cleared)
            const uint32_t gp = gen - 1u;
            if (a.nl.words[4 + (gp & 1u)] == gp) {
                const uint2 cr = reinterpret_cast<const uint2*>(a.nl.cellRec)[i];
                if (cr.y != kInvalid) {
                    uint32_t* cc = a.nl.cellCnt + (size_t)(gp & 1u) * a.nl.ncells;
                    const uint32_t sx = a.grid.sub[0], sxy = sx * a.grid.sub[1];
                    for (uint32_t q = 0; q < 8; q++)
                        if ((q & cr.y) == q) cc[cr.x + (q & 1u) + ((q >> 1) & 1u) * sx + (q >> 2) * sxy] = 0u;
                }
            }
        }
        if (NL && !use && a.mkMirror) {
            a.mkMirror[i] = mk1;
            a.pidMirror[i] = a.pidIn[i];
        }
        if (NL && !use) {
            bool stale = nfront + nback > a.nl.cap;
            a.nl.count[i] = stale ? 0u : nfront | (nback << 16);
            reinterpret_cast<float4*>(a.nl.ref)[i] = xi;
            // count this particle in every bucket within ext of it (it can be in no other
            // bucket while the lists are used); the buckets must keep their slots
            uint2 cr = make_uint2(0u, kInvalid);
            if (isfinite(xi.x) && isfinite(xi.y) && isfinite(xi.z)) {
                const GridP& gp = a.grid;
                const float e = a.nl.ext;
                uint32_t lo[3], m = 0;
                const float xs[3] = {xi.x, xi.y, xi.z};
                for (int ax = 0; ax < 3; ax++) {
                    lo[ax] = cell_coord(xs[ax] - e, gp.min[ax], gp.dim[ax], gp.sub[ax]);
                    const uint32_t hi = cell_coord(xs[ax] + e, gp.min[ax], gp.dim[ax], gp.sub[ax]);
                    if (hi > lo[ax]) m |= 1u << ax;  // ext < bucket size: at most two per axis
                }
                const uint32_t sx = gp.sub[0], sxy = sx * gp.sub[1];
                cr.x = lo[0] + sx * (lo[1] + gp.sub[1] * lo[2]);
                cr.y = m;
                uint32_t* cc = a.nl.cellCnt + (size_t)(gen & 1u) * a.nl.ncells;
                for (uint32_t q = 0; q < 8; q++) {
                    if ((q & m) != q) continue;
                    const uint32_t key = cr.x + (q & 1u) + ((q >> 1) & 1u) * sx + (q >> 2) * sxy;
                    const uint32_t c1 = atomicAdd(&cc[key], 1u) + 1u;
                    if (c1 > a.nl.capThresh) {
                        const uint32_t used = (a.bP[key + 1] - a.bP[key]) + (a.nl.rP[key + 1] - a.nl.rP[key]);
                        stale = stale || c1 > kBucketCap - used;
                    }
                }
            } else {
                stale = true;
            }
            reinterpret_cast<uint2*>(a.nl.cellRec)[i] = cr;
            if (stale) a.nl.words[2 + (gen & 1u)] = gen;
        }

        // derivative (fluid.rs:243-247): den' = sum, vel' = F/den + (0,-g,0), pos' = vel
        // (i passes through an empty asm so the epilogue's addresses are formed here, not
        // held in registers across the neighbour loops)
        uint32_t mte = mat1;
        asm volatile("" : "+v"(i), "+v"(mte));
        const DMat& Me = a.mats[mte];
        const float id1 = fdiv(1.0f, d1);
        const float4 kP = make_float4(vi.x, vi.y, vi.z, acc.den);
        const float4 kV = make_float4(acc.fx * id1, acc.fy * id1 - a.phys.g, acc.fz * id1, 0.0f);
        if (a.mode == 2 || (a.skip_elastic && elastic1)) {
            reinterpret_cast<float4*>(a.KdP)[i] = kP;
            reinterpret_cast<float4*>(a.KdV)[i] = kV;
        } else {
            float4 op;
            rk_update_particle(a.rk, a.grid, Me, i, kP, kV, haveKey, nextKey, nextBin, op);
            if (NL && a.nl.srad2 > 0.0f) {
                // two-level lists: the short list's build position and the particle's remaining
                // reach of the long list, read after a short walk and written after a build
                // or a refresh (which is where the short list was built)
                float4 r0;
                if (swalk) {
                    r0 = reinterpret_cast<const float4*>(a.nl.sref)[i];
                } else {
                    float reach = a.nl.lim;
                    if (use) {
                        const float4 rl = reinterpret_cast<const float4*>(a.nl.ref)[i];
                        const float ex = xi.x - rl.x, ey = xi.y - rl.y, ez = xi.z - rl.z;
                        reach = a.nl.lim - sqrtf(ex * ex + ey * ey + ez * ez);
                    }
                    r0 = make_float4(xi.x, xi.y, xi.z, reach);
                    reinterpret_cast<float4*>(a.nl.sref)[i] = r0;
                    a.nl.scount[i] = sfront + sback > a.nl.cap ? 0u : sfront | (sback << 16);
                }
                const float dx = op.x - r0.x, dy = op.y - r0.y, dz = op.z - r0.z;
                const float d2 = dx * dx + dy * dy + dz * dz;
                // (a particle past either reach, or a non-finite one, fails both tests)
                if (!(d2 <= a.nl.slim2)) a.nl.words[16 + (gen & 1u)] = gen;
                if (!(r0.w >= 0.0f && d2 <= r0.w * r0.w)) a.nl.words[2 + (gen & 1u)] = gen;
            } else if (NL) {
                // the next stage may keep the lists only if this particle stays within lim of
                // its position at the build
                const float4 r0 = use ? reinterpret_cast<const float4*>(a.nl.ref)[i] : X[i];
                const float dx = op.x - r0.x, dy = op.y - r0.y, dz = op.z - r0.z;
                if (!(dx * dx + dy * dy + dz * dz <= a.nl.lim2)) a.nl.words[2 + (gen & 1u)] = gen;
            }
        }
    }
    if (a.rk.keys && a.mode == 0) {
        const uint32_t rank = agg_rank(a.rk.binCount, nextBin, haveKey);
        if (haveKey) {
            a.rk.keys[i] = nextKey;
            a.rk.ranks[i] = rank;
        }
    }
}

// the list words as an aborted speculative step found them (one thread)
extern "C" __global__ void __launch_bounds__(64) k_spec_restore(SpecRestoreArgs a) {
    if (threadIdx.x != 0) return;
    if (a.snap) {
        for (int q = 0; q < 15; q++) a.words[q] = a.snap[q];
        *a.stamp = a.snap[15];
        a.stats[0] = a.snap[16];
        for (int q = 16; q < 19; q++) a.words[q] = a.snap[q + 1];
        for (uint32_t q = 0; q < a.ndec; q++) a.stampDec[q] -= 1u;
    }
    a.words[kNlAbort] = kInvalid;
}

// ------------------------------------------------------------------------------------
// Entry points of the step kernels: one scene per launch, or a batch of independent scenes
// in one launch (each block looks up its scene's argument block and its block index within
// that scene).
// ------------------------------------------------------------------------------------
template <class A>
__device__ __forceinline__ const A& batch_scene(const BatchArgs& b, uint32_t L, uint32_t& bid, uint32_t& nb) {
    const uint32_t s = b.blockScene[L];
    bid = L - b.blockFirst[s];
    nb = b.blockFirst[s + 1] - b.blockFirst[s];
    // the blocks are never written while a kernel runs: read them through the constant
    // address space, so that their fields come in with scalar loads (through a global
    // pointer the compiler must assume the kernel's own stores may change them, and loads
    // every field the hot loops use with vector loads)
    typedef const __attribute__((address_space(4))) A CA;
    typedef const __attribute__((address_space(4))) char CC;
    return *(const A*)(CA*)((CC*)(const void*)b.args + (size_t)s * b.stride);
}

#define ARP_ENTRIES(name, Args, body)                                                     \
    extern "C" __global__ void __launch_bounds__(256) name(Args a) { body(a, blockIdx.x, gridDim.x); } \
    extern "C" __global__ void __launch_bounds__(256) name##_batch(BatchArgs b) {         \
        uint32_t bid, nb;                                                                 \
        const Args& a = batch_scene<Args>(b, blockIdx.x, bid, nb);                        \
        body(a, bid, nb);                                                                 \
    }

ARP_ENTRIES(k_key, KeyArgs, key_body)
ARP_ENTRIES(k_scan, ScanArgs, scan_body)
ARP_ENTRIES(k_scatter, ScatterArgs, scatter_body)
ARP_ENTRIES(k_rowsort, RowSortArgs, rowsort_body)

// k_fluid: blocks are remapped so that each XCD processes a contiguous share of the
// cell-sorted particles (of the whole batch)
// (five waves per SIMD: measured faster than four without the few spilled registers)
// the material tables' location is chosen once per launch (uniform branch)
template <int MODE, bool DIRECT = (ARP_WALK_DIRECT != 0)>
__device__ __forceinline__ void fluid_entry(const FluidArgs& a, uint32_t bid, uint32_t nb) {
    // a list build (MODE 2) may be launched with fewer workgroups than blocks of 256
    // particles (in a batch, where it mostly returns at once): each workgroup then takes
    // every nb-th block
    const uint32_t nvb = MODE == 2 ? (a.n + 255u) / 256u : bid + 1u;
    for (uint32_t vb = bid; vb < nvb; vb += nb) {
        if (a.phys.nm <= (uint32_t)kMatLds)
            fluid_body<MODE, true, false, DIRECT>(a, vb, nb);
        else
            fluid_body<MODE, false, false, DIRECT>(a, vb, nb);
        if (MODE != 2) break;
    }
}
template <int MODE, bool DIRECT = (ARP_WALK_DIRECT != 0)>
__device__ __forceinline__ void fluid_entry_batch(const BatchArgs& b) {
    uint32_t bid, nb;
    const FluidArgs& a = batch_scene<FluidArgs>(b, xcd_block(), bid, nb);
    fluid_entry<MODE, DIRECT>(a, bid, nb);
}

// (five waves per SIMD: measured faster than four without the few spilled registers)
extern "C" __global__ void __launch_bounds__(256, 5) k_fluid(FluidArgs a) { fluid_entry<0>(a, xcd_block(), gridDim.x); }
extern "C" __global__ void __launch_bounds__(256, 5) k_fluid_batch(BatchArgs b) { fluid_entry_batch<0>(b); }
extern "C" __global__ void __launch_bounds__(256) k_fluid_rec(FluidArgs a) { fluid_entry<1>(a, xcd_block(), gridDim.x); }
// the list pair of kernels: k_fluid_nlb searches and builds, k_fluid_nl walks the lists
extern "C" __global__ void __launch_bounds__(256, 4) k_fluid_nlb(FluidArgs a) { fluid_entry<2>(a, xcd_block(), gridDim.x); }
extern "C" __global__ void __launch_bounds__(256, 4) k_fluid_nlb_batch(BatchArgs b) { fluid_entry_batch<2>(b); }
#ifndef ARP_NL_WAVES
#define ARP_NL_WAVES 4  // waves per SIMD the list walks are compiled for (the direct walk's registers)
#endif
extern "C" __global__ void __launch_bounds__(256, ARP_NL_WAVES) k_fluid_nl(FluidArgs a) { fluid_entry<3>(a, xcd_block(), gridDim.x); }
extern "C" __global__ void __launch_bounds__(256) k_fluid_nl_batch(BatchArgs b) { fluid_entry_batch<3>(b); }
// the same pair recording each stage's neighbours within h (the reverse sweep reads them)
extern "C" __global__ void __launch_bounds__(256) k_fluid_nlbr(FluidArgs a) { fluid_entry<4>(a, xcd_block(), gridDim.x); }
extern "C" __global__ void __launch_bounds__(256) k_fluid_nlr(FluidArgs a) { fluid_entry<5>(a, xcd_block(), gridDim.x); }
// the walks of scenes with one mobile and one boundary material and no elastic particles
extern "C" __global__ void __launch_bounds__(256, ARP_NL_WAVES) k_fluid_nl_f(FluidArgs a) {
    fluid_body<3, false, true>(a, xcd_block(), gridDim.x);
}
extern "C" __global__ void __launch_bounds__(256) k_fluid_nlr_f(FluidArgs a) {
    fluid_body<5, false, true>(a, xcd_block(), gridDim.x);
}
// (a batch of scenes that all have one mobile and one boundary material, each its own)
extern "C" __global__ void __launch_bounds__(256, ARP_NL_WAVES) k_fluid_nl_f_batch(BatchArgs b) {
    uint32_t bid, nb;
    const FluidArgs& a = batch_scene<FluidArgs>(b, xcd_block(), bid, nb);
    fluid_body<3, false, true>(a, bid, nb);
}
// a batched stage's list pair in one launch (scenes of one mobile and one boundary material):
// a block whose scene's stage builds searches and builds its block of particles, one whose
// stage reuses walks it; the body that does not apply returns at once.  Both bodies read the
// decision from the previous stage's words, which neither writes, so they agree.
extern "C" __global__ void __launch_bounds__(256, ARP_NL_WAVES) k_fluid_nlbw_f_batch(BatchArgs b) {
    uint32_t bid, nb;
    const FluidArgs& a = batch_scene<FluidArgs>(b, xcd_block(), bid, nb);
    if (a.phys.nm <= (uint32_t)kMatLds)
        fluid_body<2, true, false>(a, bid, nb);
    else
        fluid_body<2, false, false>(a, bid, nb);
    fluid_body<3, false, true>(a, bid, nb);
}
// the batch walks' other form, a choice of the batch autotuner: the entries within h
// buffered in LDS and drained with their positions loaded again (five waves per SIMD)
extern "C" __global__ void __launch_bounds__(256, 5) k_fluid_nl_f_batch_buf(BatchArgs b) {
    uint32_t bid, nb;
    const FluidArgs& a = batch_scene<FluidArgs>(b, xcd_block(), bid, nb);
    fluid_body<3, false, true, false>(a, bid, nb);
}
extern "C" __global__ void __launch_bounds__(256) k_fluid_nl_batch_buf(BatchArgs b) { fluid_entry_batch<3, false>(b); }

// ------------------------------------------------------------------------------------
// std430 AoS <-> internal SoA
// ------------------------------------------------------------------------------------
extern "C" __global__ void __launch_bounds__(256) k_aos_in(AosInArgs a) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= a.n) return;
    const AosParticle p = reinterpret_cast<const AosParticle*>(a.aos)[i];
    const uint32_t mk = a.mkStatic[i];
    const DMat& M = a.mats[mk & kMatMask];
    reinterpret_cast<float4*>(a.posd)[i] = make_float4(p.pos.x, p.pos.y, p.pos.z, p.den);
    reinterpret_cast<float4*>(a.velq)[i] = make_float4(p.vel.x, p.vel.y, p.vel.z, pressure_q(M, p.den));
    a.mk[i] = mk;
    a.pid[i] = i;
}

extern "C" __global__ void __launch_bounds__(256) k_aos_out(AosOutArgs a) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= a.n) return;
    const float4 p = reinterpret_cast<const float4*>(a.posd)[i];
    const float4 v = reinterpret_cast<const float4*>(a.velq)[i];
    const uint32_t id = a.pid[i];
    AosParticle o;
    o.den = p.w;
    o.mat = a.matByPid[id];
    o.solid_id = a.sidByPid[id];
    o.pad = 0;
    o.pos = make_float4(p.x, p.y, p.z, 0.0f);
    o.vel = make_float4(v.x, v.y, v.z, 0.0f);
    reinterpret_cast<AosParticle*>(a.aos)[id] = o;
}

// lin_comb_N / s_lin_comb_N (particles.rs:8-154): dst = sum_t r_t * src_t in term order
extern "C" __global__ void __launch_bounds__(256) k_lincomb(LincombArgs a) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i < a.n) {
        const AosParticle* t0 = reinterpret_cast<const AosParticle*>(a.terms[0]);
        AosParticle o = t0[i];
        float c = a.coef[0];
        float den = c * o.den;
        float4 p = make_float4(c * o.pos.x, c * o.pos.y, c * o.pos.z, c * o.pos.w);
        float4 v = make_float4(c * o.vel.x, c * o.vel.y, c * o.vel.z, c * o.vel.w);
        for (uint32_t k = 1; k < a.nterms; k++) {
            const AosParticle q = reinterpret_cast<const AosParticle*>(a.terms[k])[i];
            c = a.coef[k];
            den = den + c * q.den;
            p.x = p.x + c * q.pos.x; p.y = p.y + c * q.pos.y; p.z = p.z + c * q.pos.z; p.w = p.w + c * q.pos.w;
            v.x = v.x + c * q.vel.x; v.y = v.y + c * q.vel.y; v.z = v.z + c * q.vel.z; v.w = v.w + c * q.vel.w;
        }
        o.den = den;
        o.pos = p;
        o.vel = v;
        reinterpret_cast<AosParticle*>(a.out)[i] = o;
    }
    if (i < a.ns && a.sout) {
        const AosSolid* s0 = reinterpret_cast<const AosSolid*>(a.sterms[0]);
        AosSolid o = s0[i];
        float c = a.coef[0];
        float r[20];
        r[0] = c * o.ref_pos.x; r[1] = c * o.ref_pos.y; r[2] = c * o.ref_pos.z; r[3] = c * o.ref_pos.w;
        for (int q = 0; q < 4; q++) {
            r[4 + 4 * q] = c * o.stress[q].x; r[5 + 4 * q] = c * o.stress[q].y;
            r[6 + 4 * q] = c * o.stress[q].z; r[7 + 4 * q] = c * o.stress[q].w;
        }
        for (uint32_t k = 1; k < a.nterms; k++) {
            const AosSolid s = reinterpret_cast<const AosSolid*>(a.sterms[k])[i];
            c = a.coef[k];
            r[0] = r[0] + c * s.ref_pos.x; r[1] = r[1] + c * s.ref_pos.y;
            r[2] = r[2] + c * s.ref_pos.z; r[3] = r[3] + c * s.ref_pos.w;
            for (int q = 0; q < 4; q++) {
                r[4 + 4 * q] = r[4 + 4 * q] + c * s.stress[q].x;
                r[5 + 4 * q] = r[5 + 4 * q] + c * s.stress[q].y;
                r[6 + 4 * q] = r[6 + 4 * q] + c * s.stress[q].z;
                r[7 + 4 * q] = r[7 + 4 * q] + c * s.stress[q].w;
            }
        }
        o.ref_pos = make_float4(r[0], r[1], r[2], r[3]);
        for (int q = 0; q < 4; q++) o.stress[q] = make_float4(r[4 + 4 * q], r[5 + 4 * q], r[6 + 4 * q], r[7 + 4 * q]);
        reinterpret_cast<AosSolid*>(a.sout)[i] = o;
    }
}

// vel / vel_solid (particles.rs:194-203, 242-249)
extern "C" __global__ void __launch_bounds__(256) k_velocity(VelocityArgs a) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i < a.n) {
        const AosParticle p = reinterpret_cast<const AosParticle*>(a.in)[i];
        AosParticle o;
        o.mat = p.mat;
        o.solid_id = p.solid_id;
        o.pad = 0;
        o.den = 0.0f;
        o.pos = make_float4(a.r * p.vel.x, a.r * p.vel.y, a.r * p.vel.z, a.r * p.vel.w);
        o.vel = make_float4(0.f, 0.f, 0.f, 0.f);
        reinterpret_cast<AosParticle*>(a.out)[i] = o;
    }
    if (i < a.ns && a.sout) {
        const AosSolid s = reinterpret_cast<const AosSolid*>(a.sin)[i];
        AosSolid o;
        o.part_id = s.part_id;
        o.pad[0] = o.pad[1] = o.pad[2] = 0;
        o.ref_pos = make_float4(0.f, 0.f, 0.f, 0.f);
        for (int q = 0; q < 4; q++) o.stress[q] = make_float4(0.f, 0.f, 0.f, 0.f);
        reinterpret_cast<AosSolid*>(a.sout)[i] = o;
    }
}

// u32 fill (clears counters without hipMemsetAsync, so the chain stays graph-capturable)
extern "C" __global__ void __launch_bounds__(256) k_fill(FillArgs a) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i < a.n) a.p[i] = a.v;
}

// clear_solids (solid.rs:5-37): dest = {part_id, ref_pos = 0, stress = 0}
extern "C" __global__ void __launch_bounds__(256) k_clear_solids(ClearSolidsArgs a) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= a.ns) return;
    AosSolid o;
    o.part_id = reinterpret_cast<const AosSolid*>(a.in)[i].part_id;
    o.pad[0] = o.pad[1] = o.pad[2] = 0;
    o.ref_pos = make_float4(0.f, 0.f, 0.f, 0.f);
    for (int q = 0; q < 4; q++) o.stress[q] = make_float4(0.f, 0.f, 0.f, 0.f);
    reinterpret_cast<AosSolid*>(a.out)[i] = o;
}

// ------------------------------------------------------------------------------------
// solids: corrected-gradient strain (strain.rs:71-181) and solid forces
// (solid.rs:276-524).  Matrices are GLSL mat4, column-major m[4*col + row].
// ------------------------------------------------------------------------------------
struct M4 {
    float m[16];
};

__device__ __forceinline__ M4 m4_zero() {
    M4 r;
#pragma unroll
    for (int q = 0; q < 16; q++) r.m[q] = 0.0f;
    return r;
}
__device__ __forceinline__ M4 m4_ident() {
    M4 r = m4_zero();
    r.m[0] = r.m[5] = r.m[10] = r.m[15] = 1.0f;
    return r;
}
__device__ __forceinline__ M4 m4_mul(const M4& A, const M4& B) {
    M4 t;
#pragma unroll
    for (int c = 0; c < 4; c++)
#pragma unroll
        for (int r = 0; r < 4; r++)
            t.m[4 * c + r] = A.m[r] * B.m[4 * c] + A.m[4 + r] * B.m[4 * c + 1] + A.m[8 + r] * B.m[4 * c + 2] +
                             A.m[12 + r] * B.m[4 * c + 3];
    return t;
}
__device__ __forceinline__ M4 m4_trans(const M4& A) {
    M4 t;
#pragma unroll
    for (int c = 0; c < 4; c++)
#pragma unroll
        for (int r = 0; r < 4; r++) t.m[4 * c + r] = A.m[4 * r + c];
    return t;
}
__device__ __forceinline__ M4 m4_add(const M4& A, const M4& B) {
    M4 t;
#pragma unroll
    for (int q = 0; q < 16; q++) t.m[q] = A.m[q] + B.m[q];
    return t;
}
__device__ __forceinline__ M4 m4_scale(float s, const M4& A) {
    M4 t;
#pragma unroll
    for (int q = 0; q < 16; q++) t.m[q] = s * A.m[q];
    return t;
}
__device__ M4 m4_inv(const M4& M) {
    const float* m = M.m;
    float a00 = m[0], a01 = m[1], a02 = m[2], a03 = m[3];
    float a10 = m[4], a11 = m[5], a12 = m[6], a13 = m[7];
    float a20 = m[8], a21 = m[9], a22 = m[10], a23 = m[11];
    float a30 = m[12], a31 = m[13], a32 = m[14], a33 = m[15];
    float b00 = a00 * a11 - a01 * a10, b01 = a00 * a12 - a02 * a10;
    float b02 = a00 * a13 - a03 * a10, b03 = a01 * a12 - a02 * a11;
    float b04 = a01 * a13 - a03 * a11, b05 = a02 * a13 - a03 * a12;
    float b06 = a20 * a31 - a21 * a30, b07 = a20 * a32 - a22 * a30;
    float b08 = a20 * a33 - a23 * a30, b09 = a21 * a32 - a22 * a31;
    float b10 = a21 * a33 - a23 * a31, b11 = a22 * a33 - a23 * a32;
    float det = b00 * b11 - b01 * b10 + b02 * b09 + b03 * b08 - b04 * b07 + b05 * b06;
    float id = 1.0f / det;
    M4 t;
    t.m[0] = (a11 * b11 - a12 * b10 + a13 * b09) * id;
    t.m[1] = (a02 * b10 - a01 * b11 - a03 * b09) * id;
    t.m[2] = (a31 * b05 - a32 * b04 + a33 * b03) * id;
    t.m[3] = (a22 * b04 - a21 * b05 - a23 * b03) * id;
    t.m[4] = (a12 * b08 - a10 * b11 - a13 * b07) * id;
    t.m[5] = (a00 * b11 - a02 * b08 + a03 * b07) * id;
    t.m[6] = (a32 * b02 - a30 * b05 - a33 * b01) * id;
    t.m[7] = (a20 * b05 - a22 * b02 + a23 * b01) * id;
    t.m[8] = (a10 * b10 - a11 * b08 + a13 * b06) * id;
    t.m[9] = (a01 * b08 - a00 * b10 - a03 * b06) * id;
    t.m[10] = (a30 * b04 - a31 * b02 + a33 * b00) * id;
    t.m[11] = (a21 * b02 - a20 * b04 - a23 * b00) * id;
    t.m[12] = (a11 * b07 - a10 * b09 - a12 * b06) * id;
    t.m[13] = (a00 * b09 - a01 * b07 + a02 * b06) * id;
    t.m[14] = (a31 * b01 - a30 * b03 - a32 * b00) * id;
    t.m[15] = (a20 * b03 - a21 * b01 + a22 * b00) * id;
    return t;
}
// (loops over the dimension are unrolled with constant indices, so that matrices stay in
// registers)
__device__ __forceinline__ float m4_trace(const M4& A, uint32_t dim) {
    float t = A.m[0];
#pragma unroll
    for (uint32_t i = 1; i < 4; i++)
        if (i < dim) t += A.m[5 * i];
    return t;
}
__device__ __forceinline__ float m4_dot(const M4& x, const M4& y) {
    float s = 0.0f;
#pragma unroll
    for (int c = 0; c < 4; c++)
        s += x.m[4 * c] * y.m[4 * c] + x.m[4 * c + 1] * y.m[4 * c + 1] + x.m[4 * c + 2] * y.m[4 * c + 2] +
             x.m[4 * c + 3] * y.m[4 * c + 3];
    return s;
}
// hooke: normal*trace(strain)*I + 2*shear*strain (solid.rs:139-141)
__device__ __forceinline__ M4 m4_hooke(const M4& S, float lambda, float mu, uint32_t dim) {
    const float tr = lambda * m4_trace(S, dim), s2 = 2.0f * mu;
    M4 o;
#pragma unroll
    for (int c = 0; c < 4; c++)
#pragma unroll
        for (int r = 0; r < 4; r++) o.m[4 * c + r] = (c == r ? tr : 0.0f) + s2 * S.m[4 * c + r];
    return o;
}
__device__ M4 m4_cauchy(const M4& F, uint32_t dim) {
    M4 C = m4_mul(m4_trans(F), F);
#pragma unroll
    for (uint32_t i = 0; i < 4; i++)
        if (i >= dim) C.m[5 * i] = 1.0f;
    return C;
}
// Seth-Hill strain of order two_m/2 (solid.rs:105-168): log series, Denman-Beavers root
__device__ M4 m4_seth_hill(const M4& F, int two_m, uint32_t dim) {
    const M4 C = m4_cauchy(F, dim);
    const M4 I = m4_ident();
    if (two_m == 0) {
        M4 sum = m4_zero(), term = m4_scale(-1.0f, I), AmI;
        for (int q = 0; q < 16; q++) AmI.m[q] = -1.0f * (C.m[q] - I.m[q]);
        for (int i = 1; i <= 5; i++) {
            term = m4_mul(term, AmI);
            for (int q = 0; q < 16; q++) sum.m[q] += term.m[q] / (float)i;
        }
        return m4_scale(0.5f, sum);
    }
    const uint32_t m = (uint32_t)abs(two_m) >> 1;
    M4 Cm = I;
    for (uint32_t i = 0; i < m; i++) Cm = m4_mul(Cm, C);
    if ((abs(two_m) & 1) == 1) {
        M4 Y = C, Z = I;
        for (int i = 0; i < 15; i++) {
            const M4 iZ = m4_inv(Z), iY = m4_inv(Y);
            M4 Yn, Zn;
            for (int q = 0; q < 16; q++) {
                Yn.m[q] = 0.5f * (Y.m[q] + iZ.m[q]);
                Zn.m[q] = 0.5f * (Z.m[q] + iY.m[q]);
            }
            Y = Yn;
            Z = Zn;
        }
        Cm = m4_mul(Cm, Y);
    }
    if (two_m < 0) Cm = m4_inv(Cm);
    M4 d;
    for (int q = 0; q < 16; q++) d.m[q] = Cm.m[q] - I.m[q];
    return m4_scale(1.0f / (float)two_m, d);
}
__device__ M4 m4_seth_hill_rate(const M4& F, const M4& Fd, int two_m, uint32_t dim) {
    M4 dC = m4_mul(m4_trans(Fd), F);
    dC = m4_add(dC, m4_trans(dC));
    const M4 C = m4_cauchy(F, dim);
    if (two_m == 0) return m4_mul(m4_scale(0.5f, dC), m4_inv(C));
    const uint32_t m = (uint32_t)abs(two_m) >> 1;
    M4 sr = m4_zero();
    for (uint32_t i = 0; i < m; i++) {
        M4 term = m4_ident();
        for (uint32_t j = 0; j < m; j++) {
            if (j == i)
                term = m4_mul(term, dC);
            else
                term = m4_mul(term, C);
        }
        sr = m4_add(sr, term);
    }
    sr = m4_scale(1.0f / (float)two_m, sr);
    if (two_m < 0) {
        const M4 iC = m4_inv(C);
        return m4_mul(m4_mul(m4_scale(-1.0f, iC), sr), iC);
    }
    return sr;
}

__device__ __forceinline__ M4 ld_m4(const float* base, uint32_t s) {
    const float4* p = reinterpret_cast<const float4*>(base) + 4 * (size_t)s;
    M4 r;
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const float4 v = p[q];
        r.m[4 * q] = v.x, r.m[4 * q + 1] = v.y, r.m[4 * q + 2] = v.z, r.m[4 * q + 3] = v.w;
    }
    return r;
}
// A and F of a solid: three rows of 12 bytes (kRows3 floats per solid), so that the gathers
// of a neighbour's matrix move 36 bytes, not three padded 16-byte rows (the solid kernels are
// bound by the data their gathers return)
constexpr uint32_t kRows3 = 9;
__device__ __forceinline__ void store_rows3(float* dst, const M4& A) {
    float3* d = reinterpret_cast<float3*>(dst);
    d[0] = make_float3(A.m[0], A.m[4], A.m[8]);
    d[1] = make_float3(A.m[1], A.m[5], A.m[9]);
    d[2] = make_float3(A.m[2], A.m[6], A.m[10]);
}

__device__ __forceinline__ bool in_grid(const GridP& g, float4 p) {
    return p.x >= g.min[0] && p.x <= g.min[0] + g.dim[0] && p.y >= g.min[1] && p.y <= g.min[1] + g.dim[1] &&
           p.z >= g.min[2] && p.z <= g.min[2] + g.dim[2];
}

// Self-contact row boxes: for each row of the current grid (one wave per row), the box of
// the displacements u = x - X of the row's solids and a flag when one of them lies outside
// the grid (its cell is clamped).  A contact pair needs |X_j - X_i| > 8h (solid.rs:426-450);
// with X = x - u, |X_j - X_i| <= |x_j - x_i| + |u_j - u_i|, and for j in the stencil of i
// (both inside the grid) |x_j - x_i| < 2 |cell|.  contact_row_clear bounds |u_j - u_i| by
// the box, so a sweep skips only rows none of whose solids can pass the test.
__device__ __forceinline__ void contact_row_boxes(const SolidArgs& a, uint32_t bid, uint32_t nb) {
    if (!a.rowBox) return;
    const uint32_t lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
    const uint32_t stampv = *a.stamp;
    const float4* X = reinterpret_cast<const float4*>(a.Xposd);
    const float4* REF = reinterpret_cast<const float4*>(a.ref);
    float4* box = reinterpret_cast<float4*>(a.rowBox);
    const float inf = __builtin_inff();
    for (uint32_t row = bid * 4u + w; row < a.grid.nrows; row += nb * 4u) {
        if (a.rowStamp[row] != stampv) continue;
        const uint32_t* ro = a.rowOff + (size_t)row * a.grid.rowstride;
        const uint32_t lo = ro[0], hi = ro[a.grid.sub[0]];
        float mn0 = inf, mn1 = inf, mn2 = inf, mx0 = -inf, mx1 = -inf, mx2 = -inf;
        uint32_t out = 0;
        for (uint32_t q = lo + lane; q < hi; q += 64u) {
            const uint32_t pe = a.perm[q];
            if (pe & kFlag) continue;
            if (!(a.mk[pe] & kMkSolid)) continue;
            const float4 X2 = REF[a.sidByPid[a.pid[pe]]];
            const float4 xj = X[pe];
            const float u0 = xj.x - X2.x, u1 = xj.y - X2.y, u2 = xj.z - X2.z;
            mn0 = fminf(mn0, u0), mn1 = fminf(mn1, u1), mn2 = fminf(mn2, u2);
            mx0 = fmaxf(mx0, u0), mx1 = fmaxf(mx1, u1), mx2 = fmaxf(mx2, u2);
            if (!in_grid(a.grid, xj)) out = 1u;
        }
        for (int o = 32; o >= 1; o >>= 1) {
            mn0 = fminf(mn0, __shfl_xor(mn0, o)), mn1 = fminf(mn1, __shfl_xor(mn1, o));
            mn2 = fminf(mn2, __shfl_xor(mn2, o));
            mx0 = fmaxf(mx0, __shfl_xor(mx0, o)), mx1 = fmaxf(mx1, __shfl_xor(mx1, o));
            mx2 = fmaxf(mx2, __shfl_xor(mx2, o));
            out |= __shfl_xor(out, o);
        }
        if (lane == 0) {
            box[2 * (size_t)row] = make_float4(mn0, mn1, mn2, __uint_as_float(out));
            box[2 * (size_t)row + 1] = make_float4(mx0, mx1, mx2, 0.0f);
        }
    }
}

// How far u_j - u_i may reach for a pair of the stencil of a solid at x (reference X) to
// stay within r_cut in the reference state: r_cut - 2 |cell| less a rounding margin, or -1
// (no row is clear) outside the grid or without boxes.
__device__ __forceinline__ float contact_reach(const SolidArgs& a, float4 x, float r_cut) {
    if (!a.rowBox || !in_grid(a.grid, x)) return -1.0f;
    const GridP& g = a.grid;
    float cell2 = 0.0f, ext = 0.0f;
    for (int d = 0; d < 3; d++) {
        const float e = g.dim[d] > 0.0f ? g.dim[d] / (float)g.sub[d] : 0.0f;
        cell2 += e * e;
        ext += fabsf(g.min[d]) + g.dim[d];
    }
    return r_cut - 2.0f * sqrtf(cell2) - (1e-5f * ext + 1e-3f * r_cut);
}

// True when no solid of the row can pass the reference-distance test against the solid of
// displacement u (see contact_row_boxes).
__device__ __forceinline__ bool contact_row_clear(const SolidArgs& a, uint32_t row, float u0, float u1, float u2,
                                                  float reach) {
    if (!(reach > 0.0f)) return false;
    const float4 lo = reinterpret_cast<const float4*>(a.rowBox)[2 * (size_t)row];
    const float4 hi = reinterpret_cast<const float4*>(a.rowBox)[2 * (size_t)row + 1];
    if (__float_as_uint(lo.w) != 0u) return false;
    if (!(lo.x <= hi.x)) return true;  // no solid in the row
    const float d0 = fmaxf(hi.x - u0, u0 - lo.x), d1 = fmaxf(hi.y - u1, u1 - lo.y), d2 = fmaxf(hi.z - u2, u2 - lo.z);
    return d0 * d0 + d1 * d1 + d2 * d2 < reach * reach;
}

// One thread per solid: F and F' from the static reference neighbours, the PK2 stress
// hooke(E(F)+eps_p) + hooke(E'), A = (F S^T) L^-1 for the elastic force, and the plastic
// flow rate (solid.rs:334-341, 497-518).
// The workgroups first compute the self-contact row boxes of the stage.
__device__ __forceinline__ void strain_body(const SolidArgs& a, uint32_t bid, uint32_t nb) {
    contact_row_boxes(a, bid, nb);
    const uint32_t s = bid * 256u + threadIdx.x;
    if (s >= a.ns) return;
    float4* ke = reinterpret_cast<float4*>(a.keps) + 4 * (size_t)s;
    const float4 z4 = make_float4(0.f, 0.f, 0.f, 0.f);
    const uint32_t p = a.part[s];
    if (p == kInvalid) {
        if (!a.epsY)
            for (int q = 0; q < 4; q++) ke[q] = z4;
        return;
    }
    const uint32_t i = a.inv[p];
    const uint32_t mk = a.mk[i];
    if (!(mk & kMkElastic)) {
        if (!a.epsY)
            for (int q = 0; q < 4; q++) ke[q] = z4;
        return;
    }
    const DMat& M = a.mats[mk & kMatMask];
    const bool fixedEps = a.epsY && !M.plastic;
    const float4* X = reinterpret_cast<const float4*>(a.Xposd);
    const float4* XV = reinterpret_cast<const float4*>(a.Xvelq);
    const float4 xi = X[i], vi = XV[i];
    const float ks = a.kscale ? a.kscale[s] : 1.0f;
    float F[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, Fd[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};  // [row*3+col]
    const float4* G = reinterpret_cast<const float4*>(a.sG);
    const uint32_t q0 = a.sOff[s], q1 = a.sOff[s + 1];
    // ARP_ST_GROUP entries at a time: each level of the entry -> storage index -> state
    // gathers is issued for the whole group before the next (a tail entry reloads the last
    // one, unused)
    for (uint32_t qb = q0; qb < q1; qb += ARP_ST_GROUP) {
        uint32_t jj[ARP_ST_GROUP];
        float4 gq[ARP_ST_GROUP], xq[ARP_ST_GROUP], vq[ARP_ST_GROUP];
#pragma unroll
        for (int k = 0; k < ARP_ST_GROUP; k++) {
            const uint32_t q = min(qb + (uint32_t)k, q1 - 1u);
            jj[k] = a.sPid[q];
            gq[k] = G[q];
        }
#pragma unroll
        for (int k = 0; k < ARP_ST_GROUP; k++) jj[k] = a.inv[jj[k]];
#pragma unroll
        for (int k = 0; k < ARP_ST_GROUP; k++) {
            xq[k] = X[jj[k]];
            vq[k] = XV[jj[k]];
        }
#pragma unroll
        for (int k = 0; k < ARP_ST_GROUP; k++) {
            if (qb + (uint32_t)k >= q1) break;
            const float4 g = gq[k], xj = xq[k], vj = vq[k];
            const float dx[3] = {xj.x - xi.x, xj.y - xi.y, xj.z - xi.z};
            const float dv[3] = {vj.x - vi.x, vj.y - vi.y, vj.z - vi.z};
            const float gg[3] = {g.x, g.y, g.z};
#pragma unroll
            for (int r = 0; r < 3; r++)
#pragma unroll
                for (int c = 0; c < 3; c++) {
                    F[3 * r + c] += g.w * (dx[r] * gg[c]);
                    Fd[3 * r + c] += g.w * (dv[r] * gg[c]);
                }
        }
    }
    M4 F4 = m4_zero(), Fd4 = m4_zero();
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) {
            F4.m[4 * c + r] = F[3 * r + c];
            Fd4.m[4 * c + r] = Fd[3 * r + c];
        }
    const M4 eps = ld_m4(fixedEps ? a.epsY : a.Xeps, s);
    const uint32_t dim = a.grid.dim_n;
    const M4 E = m4_add(m4_seth_hill(F4, M.strain_order, dim), eps);
    const M4 Ed = m4_seth_hill_rate(F4, Fd4, M.strain_order, dim);
    const M4 S = m4_add(m4_hooke(E, ks * M.lambda, ks * M.mu, dim), m4_hooke(Ed, M.lambda_d, M.mu_d, dim));
    const M4 P = m4_mul(F4, m4_trans(S));
    const M4 A = m4_mul(P, ld_m4(a.Linv, s));
    store_rows3(a.A + kRows3 * (size_t)s, A);
    store_rows3(a.Fm + kRows3 * (size_t)s, F4);
    if (a.Fd) store_rows3(a.Fd + kRows3 * (size_t)s, Fd4);  // (a recorded stage: for the reverse pass)
    M4 rate = m4_zero();
    if (M.plastic) {
        // eq_plastic_strain_sq carries the integer factor (dim-1)/dim == 0 (solid.rs:212-215)
        M4 dev = eps;
        const float tr = m4_trace(eps, dim);
        for (int c = 0; c < 4; c++) dev.m[5 * c] = eps.m[5 * c] - tr / (float)dim;
        const float eq = (float)((dim - 1) / dim) * m4_dot(dev, dev);
        float sn = m4_dot(S, S);
        const float ys = (M.yield + M.hard * powf(eq, M.hard_exp / 2.0f)) * (1.0f + M.kin_hard * logf(1.0f));
        if (ys * ys <= sn) {
            sn = sqrtf(sn);
            const float sc = (sn - ys) / M.relax;
            for (int q = 0; q < 16; q++) rate.m[q] = sc * (S.m[q] / sn);
        }
    }
    if (!fixedEps)
        for (int q = 0; q < 4; q++) ke[q] = make_float4(rate.m[4 * q], rate.m[4 * q + 1], rate.m[4 * q + 2], rate.m[4 * q + 3]);
}

ARP_ENTRIES(k_strain, SolidArgs, strain_body)

__device__ __forceinline__ float3 rows3_mul(const float4* R, float x, float y, float z) {
    return make_float3(R[0].x * x + R[0].y * y + R[0].z * z, R[1].x * x + R[1].y * y + R[1].z * z,
                       R[2].x * x + R[2].y * y + R[2].z * z);
}

// One thread per solid: elastic force over same-material reference neighbours, hourglass
// restoring force for those inside the current 27-cell stencil, self contact of
// reference-distant pairs (solid.rs:343-451); then dest + dest2 (shaders/mod.rs:79) and
// the stage epilogue for the particle and its plastic strain.
__device__ __forceinline__ void solid_body(const SolidArgs& a, uint32_t bid, uint32_t) {
    const uint32_t s = bid * 256u + threadIdx.x;
    const float4* X = reinterpret_cast<const float4*>(a.Xposd);
    const float4* XV = reinterpret_cast<const float4*>(a.Xvelq);
    const float4* REF = reinterpret_cast<const float4*>(a.ref);
    bool haveKey = false;
    uint32_t i = 0, nextKey = kInvalid, nextBin = 0;
    const float h = a.phys.h, EPS = 0.1f;
    uint32_t p = kInvalid, mk = 0;
    if (s < a.ns) {
        p = a.part[s];
        if (p != kInvalid) {
            i = a.inv[p];
            mk = a.mk[i];
        }
    }
    if (s < a.ns && p != kInvalid && (mk & kMkElastic)) {
        const uint32_t mat = mk & kMatMask;
        const DMat& M = a.mats[mat];
        const float4 xi = X[i], vi = XV[i];
        const float m1 = M.mass, d1 = xi.w, V1 = m1 / d1;
        const float4 Xs = REF[s];
        const float3* A1 = reinterpret_cast<const float3*>(a.A) + 3 * (size_t)s;
        const float3* F1 = reinterpret_cast<const float3*>(a.Fm) + 3 * (size_t)s;
        const bool valid_i = a.sflag[i] == 0u;
        uint32_t ci[3] = {0, 0, 0};
        cell_of(a.grid, xi.x, xi.y, xi.z, ci[0], ci[1], ci[2]);
        float fx = 0.f, fy = 0.f, fz = 0.f;
        const float4* EG = reinterpret_cast<const float4*>(a.eG);
        const float3* AA = reinterpret_cast<const float3*>(a.A);
        const float3* FF = reinterpret_cast<const float3*>(a.Fm);
        const uint32_t q0 = a.eOff[s], q1 = a.eOff[s + 1];
        // ARP_EL_GROUP entries at a time: each level of the gathers (entry -> storage index
        // -> state) is issued for the whole group before the next; the group's pairs are then
        // summed in entry order (a tail entry reloads the last one, unused)
        for (uint32_t qb = q0; qb < q1; qb += ARP_EL_GROUP) {
            uint32_t sq[ARP_EL_GROUP], jq[ARP_EL_GROUP];
            float4 gq[ARP_EL_GROUP], xq[ARP_EL_GROUP];
            float3 Aq[ARP_EL_GROUP][3];
#pragma unroll
            for (int k = 0; k < ARP_EL_GROUP; k++) {
                const uint32_t q = min(qb + (uint32_t)k, q1 - 1u);
                sq[k] = a.eSid[q];
                jq[k] = a.ePid[q];
                gq[k] = EG[q];
            }
#pragma unroll
            for (int k = 0; k < ARP_EL_GROUP; k++) {
                jq[k] = a.inv[jq[k]];
#pragma unroll
                for (int r = 0; r < 3; r++) Aq[k][r] = AA[3 * (size_t)sq[k] + r];
            }
#pragma unroll
            for (int k = 0; k < ARP_EL_GROUP; k++) xq[k] = X[jq[k]];
#pragma unroll
            for (int k = 0; k < ARP_EL_GROUP; k++) {
                if (qb + (uint32_t)k >= q1) break;
                const uint32_t s2 = sq[k], j = jq[k];
                const float4 g = gq[k], xj = xq[k];
                const float V2 = m1 / xj.w;
                const float c = -V1 * V2;
                float4 Asum[3];
                for (int r = 0; r < 3; r++)
                    Asum[r] = make_float4(c * (A1[r].x + Aq[k][r].x), c * (A1[r].y + Aq[k][r].y),
                                          c * (A1[r].z + Aq[k][r].z), 0.f);
                const float3 el = rows3_mul(Asum, g.x, g.y, g.z);
                fx += el.x;
                fy += el.y;
                if (a.grid.dim_n > 2) fz += el.z;
                // hourglass (solid.rs:432-443): same material, W(dX) != 0, inside the stencil
                if (g.w != 0.0f && valid_i && a.sflag[j] == 0u) {
                    uint32_t cj[3] = {0, 0, 0};
                    cell_of(a.grid, xj.x, xj.y, xj.z, cj[0], cj[1], cj[2]);
                    bool adj = true;
                    for (int d = 0; d < 3; d++) adj &= (ci[d] + 1 >= cj[d]) && (cj[d] + 1 >= ci[d]);
                    if (adj) {
                        const float4 X2 = REF[s2];
                        const float dX[3] = {X2.x - Xs.x, X2.y - Xs.y, X2.z - Xs.z};
                        const float dx[3] = {xj.x - xi.x, xj.y - xi.y, xj.z - xi.z};
                        const float l = sqrtf(dx[0] * dx[0] + dx[1] * dx[1] + dx[2] * dx[2]);
                        const float3* F2 = FF + 3 * (size_t)s2;
                        float err[3];
                        for (int r = 0; r < 3; r++) {
                            const float4 fa = make_float4(0.5f * (F1[r].x + F2[r].x), 0.5f * (F1[r].y + F2[r].y),
                                                          0.5f * (F1[r].z + F2[r].z), 0.f);
                            err[r] = fa.x * dX[0] + fa.y * dX[1] + fa.z * dX[2] - dx[r];
                        }
                        const float m2 = m1, d2 = xj.w;
                        const float sc = (100000000.0f * m1 * m2 / (d1 * d2)) *
                                         (0.5f * (err[0] * dx[0] + err[1] * dx[1] + err[2] * dx[2]) / (l + 10.0f * EPS * h)) *
                                         (1.0f / (l * l + EPS * h * h)) * g.w;
                        fx -= sc * dx[0] / (l + EPS * h);
                        fy -= sc * dx[1] / (l + EPS * h);
                        fz -= sc * dx[2] / (l + EPS * h);
                    }
                }
            }
        }
        // self contact: same-material pairs further than 8h apart in the reference state but
        // inside the current stencil and closer than 8h (solid.rs:426-450)
        if (!M.plastic && valid_i) {
            const float r_cut = 8.0f * h, r_geom = h * h / r_cut;
            const uint32_t sx = a.grid.sub[0], sy = a.grid.sub[1], sz = a.grid.sub[2];
            const uint32_t xl = ci[0] > 0 ? ci[0] - 1 : 0, xh = min(ci[0] + 1, sx - 1);
            const float reach = contact_reach(a, xi, r_cut);
            const float u0 = xi.x - Xs.x, u1 = xi.y - Xs.y, u2 = xi.z - Xs.z;
            for (int dz = -1; dz <= 1; dz++) {
                const int nz = (int)ci[2] + dz;
                if (nz < 0 || nz >= (int)sz) continue;
                for (int dy = -1; dy <= 1; dy++) {
                    const int ny = (int)ci[1] + dy;
                    if (ny < 0 || ny >= (int)sy) continue;
                    const uint32_t row = (uint32_t)ny + sy * (uint32_t)nz;
                    if (a.rowStamp[row] != *a.stamp) continue;
                    if (contact_row_clear(a, row, u0, u1, u2, reach)) continue;
                    const uint32_t* ro = a.rowOff + (size_t)row * a.grid.rowstride;
                    for (uint32_t q = ro[xl]; q < ro[xh + 1]; q++) {
                        const uint32_t pe = a.perm[q];
                        if (pe & kFlag) continue;
                        const uint32_t mk2 = a.mk[pe];
                        if ((mk2 & kMatMask) != mat || !(mk2 & kMkSolid)) continue;
                        const uint32_t s2 = a.sidByPid[a.pid[pe]];
                        const float4 X2 = REF[s2];
                        const float dX0 = X2.x - Xs.x, dX1 = X2.y - Xs.y, dX2 = X2.z - Xs.z;
                        if (!(dX0 * dX0 + dX1 * dX1 + dX2 * dX2 > r_cut * r_cut)) continue;
                        const float4 xj = X[pe];
                        const float rx = xj.x - xi.x, ry = xj.y - xi.y, rz = xj.z - xi.z;
                        const float rr = rx * rx + ry * ry + rz * rz;
                        if (!(rr < r_cut * r_cut)) continue;
                        const float l = sqrtf(rr);
                        const float dr = fmaxf(r_cut - l, 0.0f);
                        const float cs = -5000.0f * sqrtf(dr * r_geom);
                        fx += cs * rx / l;
                        fy += cs * ry / l;
                        fz += cs * rz / l;
                    }
                }
            }
        }
        // dest2 (solid.rs:492-494) + dest (fluid part)
        const float4 kf = reinterpret_cast<const float4*>(a.KFp)[i];
        const float4 kfv = reinterpret_cast<const float4*>(a.KFv)[i];
        const float4 kP = make_float4(kf.x + vi.x, kf.y + vi.y, kf.z + vi.z, kf.w + 0.0f);
        const float4 kV = make_float4(kfv.x + fx / d1, kfv.y + (fy / d1 - a.phys.g), kfv.z + fz / d1, 0.0f);
        if (a.mode == 2) {
            reinterpret_cast<float4*>(a.KdP)[i] = kP;
            reinterpret_cast<float4*>(a.KdV)[i] = kV;
        } else {
            float4 op;
            rk_update_particle(a.rk, a.grid, M, i, kP, kV, haveKey, nextKey, nextBin, op);
            if (!(a.epsY && !M.plastic)) rk_update_eps(a.rk, s, a.keps + 16 * (size_t)s);
        }
    } else if (s < a.ns && a.mode == 0 && !a.epsY) {
        // solids of non-elastic materials keep their plastic strain (zero rate)
        float z[16];
        for (int q = 0; q < 16; q++) z[q] = 0.0f;
        rk_update_eps(a.rk, s, z);
    }
    if (a.rk.keys && a.mode == 0) {
        const uint32_t rank = agg_rank(a.rk.binCount, nextBin, haveKey);
        if (haveKey) {
            a.rk.keys[i] = nextKey;
            a.rk.ranks[i] = rank;
        }
    }
}

ARP_ENTRIES(k_solid, SolidArgs, solid_body)

extern "C" __global__ void __launch_bounds__(256) k_eps_in(EpsInArgs a) {
    const uint32_t s = blockIdx.x * 256u + threadIdx.x;
    if (s >= a.ns) return;
    const AosSolid so = reinterpret_cast<const AosSolid*>(a.aos)[a.order[s]];
    float4* e = reinterpret_cast<float4*>(a.eps + 16 * (size_t)s);
    for (int q = 0; q < 4; q++) e[q] = so.stress[q];
}

extern "C" __global__ void __launch_bounds__(256) k_eps_out(EpsOutArgs a) {
    const uint32_t s = blockIdx.x * 256u + threadIdx.x;
    if (s >= a.ns) return;
    AosSolid o;
    o.part_id = a.part[s];
    o.pad[0] = o.pad[1] = o.pad[2] = 0;
    o.ref_pos = a.ref ? reinterpret_cast<const float4*>(a.ref)[s] : make_float4(0.f, 0.f, 0.f, 0.f);
    const float4* e = reinterpret_cast<const float4*>(a.eps + 16 * (size_t)s);
    for (int q = 0; q < 4; q++) o.stress[q] = e[q];
    reinterpret_cast<AosSolid*>(a.aos)[a.order[s]] = o;
}

// ------------------------------------------------------------------------------------
// adjoint of the fluid stage (reverse mode of fluid.rs:153-247)
// ------------------------------------------------------------------------------------

// d/dL of contact_potential(eq, L, strength, r0) (fluid.rs:54-67)
__device__ __forceinline__ float contact_potential_d(uint32_t eq, float r, float strength, float r0) {
    switch (eq) {
        case 2: return -strength;
        case 3: {
            const float f = pow7(r0 * frcp(r));
            return (strength * (1.0f / 7.0f)) * (f - 1.0f) - strength * f;
        }
        case 4: {
            float f = r0 / r;
            f *= f * f;
            f *= f;
            return strength * (2.0f * f - 2.0f) * (-6.0f * f / r);
        }
        case 5: {
            const float x = r * frcp(r0), x2 = x * x;
            return strength * (x2 * x2 * x2);
        }
        default: return 0.0f;
    }
}

// d q / d rho with q = p(rho) / rho^2
__device__ __forceinline__ float pressure_q_d(const DMat& m, float den) {
    float p = eos_pressure(m.eq, den, m.c, m.target_den), dp = 0.0f;
    if (m.eq == 2) dp = m.c * m.c;
    if (m.eq == 3) {
        const float x = den * frcp(m.target_den), x2 = x * x;
        dp = m.c * m.c * (x2 * x2 * x2);
    }
    return dp / (den * den) - 2.0f * p / (den * den * den);
}

struct PairSide {
    float mi, mj, rhoi, rhoj, qi, qj, dqi, dqj, fi, fj, bfi, bfj;
    bool eli, elj, bnd, same;
    uint32_t mati, matj;
};

struct PairVjp {
    float gr[3], gv[3], gri, grj;
    float F[3];
    float dbulk, dvisc;
};

// Gradient of a.(pair force) + wd*(pair density rate) for target i, neighbour j with
// r = x_j - x_i, v = v_j - v_i.  Zero outside the kernel support.
__device__ void pair_vjp(const PairSide& P, const DInter& it, float rx, float ry, float rz, float vx, float vy,
                         float vz, float ax, float ay, float az, float wd, float hh, float kn, float eps_hh,
                         PairVjp& o) {
    for (int c = 0; c < 3; c++) o.gr[c] = o.gv[c] = o.F[c] = 0.0f;
    o.gri = o.grj = o.dbulk = o.dvisc = 0.0f;
    const float rr = rx * rx + ry * ry + rz * rz;
    const float d = hh - rr;
    if (!(d > 0.0f)) return;
    const float gs = -6.0f * kn * d * d, dgs = 12.0f * kn * d;  // dgs = d gs / d rr
    const float r[3] = {rx, ry, rz}, v[3] = {vx, vy, vz}, A[3] = {ax, ay, az};
    const float ar = ax * rx + ay * ry + az * rz, vr = vx * rx + vy * ry + vz * rz, av = ax * vx + ay * vy + az * vz;
    const float C = P.mi * P.mj;
    if (!P.bnd) {  // density rate: wd * mj * gs * (v.r)
        const float c1 = wd * P.mj;
        for (int c = 0; c < 3; c++) {
            o.gv[c] += c1 * gs * r[c];
            o.gr[c] += c1 * (gs * v[c] + 2.0f * dgs * vr * r[c]);
        }
    }
    {  // pressure: C (qi + qj) grad W
        const float Q = P.qi + P.qj;
        for (int c = 0; c < 3; c++) {
            o.F[c] += C * Q * gs * r[c];
            o.gr[c] += C * Q * (gs * A[c] + 2.0f * dgs * ar * r[c]);
        }
        o.gri += C * P.dqi * gs * ar;
        o.grj += C * P.dqj * gs * ar;
    }
    if (it.potential) {  // contact (radius, strength swapped as in fluid.rs:177) + friction
        const float L = sqrtf(rr);
        const float cp = contact_potential(it.potential, L, it.radius, it.strength);
        if (cp != 0.0f) {
            const float dcp = contact_potential_d(it.potential, L, it.radius, it.strength);
            const float irj = frcp(P.rhoj);
            const float cs = P.mj * cp * irj;
            const float iL = frcp(L);
            for (int c = 0; c < 3; c++) {
                o.F[c] += cs * gs * r[c];
                o.gr[c] += (P.mj * irj) * (dcp * r[c] * iL * gs * ar + cp * (gs * A[c] + 2.0f * dgs * ar * r[c]));
            }
            o.grj += -cs * gs * ar * irj;
            if ((P.eli || P.elj || P.bnd) && P.mati != P.matj && it.friction != 0.0f) {
                float t[3];
                const float irr = frcp(rr);
                for (int c = 0; c < 3; c++) t[c] = v[c] - r[c] * vr * irr;
                const float tl = sqrtf(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]);
                if (tl > 0.0f && isfinite(tl)) {
                    const float mu = it.friction;
                    const float c0 = cs * gs;
                    const float N = fabsf(c0) * L;
                    float th[3];
                    const float itl = frcp(tl);
                    for (int c = 0; c < 3; c++) th[c] = t[c] * itl;
                    const float at = A[0] * th[0] + A[1] * th[1] + A[2] * th[2];
                    float ut[3];
                    for (int c = 0; c < 3; c++) {
                        o.F[c] += mu * N * th[c];
                        ut[c] = mu * N * (A[c] - th[c] * at) * itl;
                    }
                    const float rut = r[0] * ut[0] + r[1] * ut[1] + r[2] * ut[2];
                    const float sg = c0 >= 0.0f ? 1.0f : -1.0f;
                    for (int c = 0; c < 3; c++) {
                        o.gv[c] += ut[c] - r[c] * rut * irr;
                        o.gr[c] += -(vr * irr * ut[c] + v[c] * rut * irr - 2.0f * vr * rut * r[c] * (irr * irr));
                        const float dc0 = (P.mj * irj) * (dcp * r[c] * iL * gs + cp * 2.0f * dgs * r[c]);
                        o.gr[c] += mu * at * sg * (L * dc0 + c0 * r[c] * iL);
                    }
                    o.grj += mu * at * (-N * irj);
                }
            }
        }
    }
    if (!P.bnd) {
        const float D = rr + eps_hh, iD = frcp(D), iRR = frcp(P.rhoi * P.rhoj);
        if (!P.eli && P.fi + P.fj != 0.0f) {  // viscosity: -C (fi+fj) gs rr / (rhoi rhoj D) v
            const float u = gs * rr * iD;
            const float du = (dgs * rr + gs) * iD - gs * rr * (iD * iD);
            const float K4 = C * (P.fi + P.fj) * iRR;
            const float s4 = K4 * u;
            for (int c = 0; c < 3; c++) {
                o.F[c] -= s4 * v[c];
                o.gv[c] -= s4 * A[c];
                o.gr[c] -= av * K4 * du * 2.0f * r[c];
            }
            o.gri += av * s4 * frcp(P.rhoi);
            o.grj += av * s4 * frcp(P.rhoj);
        }
        if (!P.eli) o.dvisc += -av * C * (gs * rr * iD) * iRR;
        // artificial viscosity: -C ff (v.r)/((rhoi+rhoj) D) grad W
        const float ff = P.same ? P.bfi + P.bfj : it.dampening;
        const float iS = frcp(P.rhoi + P.rhoj);
        const float E0 = -C * iS;
        const float E = E0 * ff;
        const float h5 = gs * iD, dh5 = dgs * iD - gs * (iD * iD);
        const float phi = E * vr * ar * h5;
        for (int c = 0; c < 3; c++) {
            o.F[c] += E * vr * h5 * r[c];
            o.gv[c] += E * ar * h5 * r[c];
            o.gr[c] += E * (ar * h5 * v[c] + vr * h5 * A[c] + vr * ar * dh5 * 2.0f * r[c]);
        }
        o.gri += -phi * iS;
        o.grj += -phi * iS;
        if (P.same) o.dbulk += E0 * vr * ar * h5;
    }
}

// One neighbour's data as the reverse stage reads it
struct AdjIn {
    float4 xj, vj;
    float wvx, wvy, wvz, wd;  // the neighbour's W: velocity part and density part
    uint32_t mk2;
};

// Both sides of a pair of non-elastic mobile particles of one material at once: the
// cotangent of x_i, v_i, rho_i from a_i.(force on i) + wd_i (density rate of i) and from
// a_j.(force on j) + wd_j (density rate of j).  The pressure, viscosity and artificial
// viscosity forces are antisymmetric under i <-> j, so their two contractions are one
// contraction with a_i - a_j; the density rates differ only by the mass factor; the
// contact forces differ by (m/rho) of the other side.  r = x_j - x_i, v = v_j - v_i.
// Returns the amounts to subtract from G (x_i) and Gv (v_i), to add to Gr (rho_i), the
// force on i (F) and the side-i parameter partials (as pair_vjp's side a).
struct PairSym {
    float gr[3], gv[3], gri, F[3], dbulk, dvisc;
};
__device__ __forceinline__ void pair_vjp_sym(float mi, float mj, float rhoi, float rhoj, float qi, float qj, float dqi,
                                             float f, float bf, const DInter& it, float rx, float ry, float rz,
                                             float vx, float vy, float vz, float aix, float aiy, float aiz,
                                             float ajx, float ajy, float ajz, float wdi, float wdj, float hh,
                                             float kn, float eps_hh, PairSym& o) {
    for (int c = 0; c < 3; c++) o.gr[c] = o.gv[c] = o.F[c] = 0.0f;
    o.gri = o.dbulk = o.dvisc = 0.0f;
    const float rr = rx * rx + ry * ry + rz * rz;
    const float d = hh - rr;
    if (!(d > 0.0f)) return;
    const float gs = -6.0f * kn * d * d, dgs = 12.0f * kn * d;
    const float r[3] = {rx, ry, rz}, v[3] = {vx, vy, vz};
    const float Ae[3] = {aix - ajx, aiy - ajy, aiz - ajz};
    const float vr = vx * rx + vy * ry + vz * rz;
    const float are = Ae[0] * rx + Ae[1] * ry + Ae[2] * rz;
    const float ari = aix * rx + aiy * ry + aiz * rz;
    const float ave = Ae[0] * vx + Ae[1] * vy + Ae[2] * vz;
    const float avi = aix * vx + aiy * vy + aiz * vz;
    const float C = mi * mj;
    // Every term lies along r, v, Ae (or the contact's Ac): scalar coefficients first, then
    // three products per direction.  Terms:
    //   density rates (wd_i m_j + wd_j m_i) gs (v.r):  gv += c1 gs r;  gr += c1 (gs v + 2 dgs (v.r) r)
    //   pressure C (qi + qj) grad W:  F += CQ gs r;  gr += CQ (gs Ae + 2 dgs (Ae.r) r);  gri += C dqi gs (Ae.r)
    //   artificial viscosity -C bf (v.r) / ((rhoi + rhoj) D) grad W, h5 = gs / D:
    //     F += E (v.r) h5 r;  gv += E (Ae.r) h5 r;
    //     gr += E ((Ae.r) h5 v + (v.r) h5 Ae + 2 (v.r)(Ae.r) dh5 r);  gri -= E (v.r)(Ae.r) h5 / (rhoi + rhoj)
    //   viscosity -C f gs rr / (rhoi rhoj D) v:  F -= s4 v;  gv -= s4 Ae;  gr -= 2 (Ae.v) K4 du r
    //   contact (m_j / rho_j) on side i, (m_i / rho_i) on side j:
    //     F += sa cp gs r;  gr += (dcp gs (Ac.r) / L + 2 cp dgs (Ac.r)) r + cp gs Ac
    const float c1 = wdi * mj + wdj * mi;
    const float CQ = C * (qi + qj);
    const float D = rr + eps_hh, iD = frcp(D), iRR = frcp(rhoi * rhoj);
    const float iS = frcp(rhoi + rhoj);
    const float E0 = -C * iS;
    const float E = E0 * bf;
    const float h5 = gs * iD, dh5 = dgs * iD - gs * (iD * iD);
    const float Evh = E * vr * h5, Eah = E * are * h5;
    float Fr = CQ * gs + Evh;  // F along r
    const float Ra = Fr;       // gr along Ae
    const float Rv = c1 * gs + Eah;  // gr along v, gv along r
    float Rr = 2.0f * dgs * (c1 * vr + CQ * are) + 2.0f * E * vr * are * dh5;  // gr along r
    float Fv = 0.0f, Ga = 0.0f;  // F along v, gv along Ae (viscosity)
    float gri = C * dqi * gs * are - Evh * are * iS;
    float cA = 0.0f, Ac[3] = {0.0f, 0.0f, 0.0f};  // gr along the contact's Ac
    if (it.potential) {
        const float L = sqrtf(rr);
        const float cp = contact_potential(it.potential, L, it.radius, it.strength);
        if (cp != 0.0f) {
            const float dcp = contact_potential_d(it.potential, L, it.radius, it.strength);
            const float irj = frcp(rhoj), iri = frcp(rhoi), iL = frcp(L);
            const float sa = mj * irj, sb = mi * iri;
            Ac[0] = sa * aix - sb * ajx;
            Ac[1] = sa * aiy - sb * ajy;
            Ac[2] = sa * aiz - sb * ajz;
            const float arc = Ac[0] * rx + Ac[1] * ry + Ac[2] * rz;
            const float arj = ajx * rx + ajy * ry + ajz * rz;
            Fr += sa * cp * gs;
            Rr += dcp * iL * gs * arc + 2.0f * cp * dgs * arc;
            cA = cp * gs;
            gri += sb * cp * gs * arj * iri;
        }
    }
    if (f != 0.0f) {
        const float u = gs * rr * iD;
        const float du = (dgs * rr + gs) * iD - gs * rr * (iD * iD);
        const float K4 = C * f * iRR;
        const float s4 = K4 * u;
        Fv = -s4;
        Ga = -s4;
        Rr -= 2.0f * ave * K4 * du;
        gri += ave * s4 * frcp(rhoi);
    }
    for (int c = 0; c < 3; c++) {
        o.F[c] = Fr * r[c] + Fv * v[c];
        o.gv[c] = Rv * r[c] + Ga * Ae[c];
        o.gr[c] = Rr * r[c] + Rv * v[c] + Ra * Ae[c] + cA * Ac[c];
    }
    o.gri = gri;
    o.dvisc = -avi * C * (h5 * rr) * iRR;
    o.dbulk = E0 * vr * ari * h5;
}

// REC: the neighbours come from the forward stage's record instead of a new sweep.
// FAST: every mobile particle has material a.fastMat and there are at most kMatLds
// materials: the tables are read from LDS and the parameter partials are two scalars per
// thread (the general form keeps a partial per material, indexed per pair).
// FLUID (with FAST): no elastic particle in the scene, so that every mobile pair takes the
// two-sided form and the one-sided form serves boundary pairs only.
template <bool REC, bool FAST, bool FLUID = false>
__device__ __forceinline__ void fluid_adj_body(const AdjArgs& a) {
    uint32_t* list = s_list;
    float4* smat = s_mat;
    DInter* sint = s_inter;
    __shared__ float spg[4][2 * kMatLds];  // per wave, summed in a fixed order
    const uint32_t tid = threadIdx.x;
    const uint32_t nm = a.phys.nm;
    for (uint32_t m = tid; m < 4 * 2 * kMatLds; m += 256) spg[m / (2 * kMatLds)][m % (2 * kMatLds)] = 0.0f;
    if (FAST) {
        for (uint32_t m = tid; m < nm; m += 256)
            smat[m] = make_float4(a.mats[m].mass, a.mats[m].visc, a.mats[m].bulk_visc, 0.f);
        for (uint32_t m = tid; m < nm * nm; m += 256) sint[m] = a.inter[m];
    }
    __syncthreads();
    const uint32_t t = xcd_block() * 256u + tid;
    const float4* X = reinterpret_cast<const float4*>(a.Xposd);
    const float4* XV = reinterpret_cast<const float4*>(a.Xvelq);
    const float4* B = reinterpret_cast<const float4*>(a.bposd);
    const float4* BV = reinterpret_cast<const float4*>(a.bvelq);
    const float4* WP = reinterpret_cast<const float4*>(a.Wp);
    const float4* WV = reinterpret_cast<const float4*>(a.Wv);
    const float h = a.phys.h, hh = h * h, kn = a.phys.k;
    constexpr int NP = FAST ? 1 : kMatLds;
    float pb[NP], pv[NP];  // per-thread parameter partials (FAST: of a.fastMat only)
    for (int m = 0; m < NP; m++) pb[m] = pv[m] = 0.0f;
    if (t < a.n) {
        uint32_t i;
        bool noNb;
        if (a.stageUse && *a.stageUse) {  // the forward stage walked kept lists
            i = t;
            noNb = false;
        } else if (a.identity) {
            i = t;
            noNb = a.sflag[t] != 0u;
        } else {
            const uint32_t pe = a.perm[t];
            i = pe & ~kFlag;
            noNb = (pe & kFlag) != 0;
        }
        const float4 xi = X[i], vi = XV[i];
        const uint32_t mk1 = a.mk[i];
        const uint32_t mat1 = mk1 & kMatMask;
        const DMat& M1 = a.mats[mat1];
        const float rho1 = xi.w;
        const float ws = a.wscale;
        float4 wp = WP[i], wv = WV[i];
        wp = make_float4(ws * wp.x, ws * wp.y, ws * wp.z, ws * wp.w);
        wv = make_float4(ws * wv.x, ws * wv.y, ws * wv.z, 0.0f);
        const float ir1 = frcp(rho1);
        const float a1x = wv.x * ir1, a1y = wv.y * ir1, a1z = wv.z * ir1;
        const float dq1 = pressure_q_d(M1, rho1);
        float G[3] = {0, 0, 0}, Gv[3] = {0, 0, 0}, Gr = 0.0f, F[3] = {0, 0, 0};
        if (!noNb) {
            // the recorded / found neighbours; the next one's loads are issued before the
            // current pair's arithmetic, into the other of two register sets
            auto ld = [&](uint32_t e) {
                AdjIn in;
                const bool bnd = (e & kBnd) != 0;
                const uint32_t j = e & ~kBnd;
                in.xj = bnd ? B[j] : X[j];
                in.vj = bnd ? BV[j] : XV[j];
                // (FLUID with FAST: every mobile particle is of material fastMat)
                in.mk2 = bnd ? a.bmk[j] : ((FLUID && FAST) ? a.fastMat : a.mk[j]);
                const float4 wvj = bnd ? make_float4(0.f, 0.f, 0.f, 0.f) : WV[j];
                in.wvx = wvj.x;
                in.wvy = wvj.y;
                in.wvz = wvj.z;
                in.wd = bnd ? 0.0f : reinterpret_cast<const float*>(WP)[4 * (size_t)j + 3];
                return in;
            };
            auto proc = [&](uint32_t e, const AdjIn& in) {
                const bool bnd = (e & kBnd) != 0;
                const float4 xj = in.xj, vj = in.vj;
                const uint32_t mk2 = in.mk2;
                const uint32_t mat2 = mk2 & kMatMask;
                float m2, f2, bf2;
                DInter itA, itB;
                if constexpr (FAST) {
                    const float4 mm = smat[mat2];
                    m2 = mm.x;
                    f2 = mm.y;
                    bf2 = mm.z;
                    itA = sint[mat1 * nm + mat2];
                    itB = sint[mat2 * nm + mat1];
                } else {
                    const DMat& M2 = a.mats[mat2];
                    m2 = M2.mass;
                    f2 = M2.visc;
                    bf2 = M2.bulk_visc;
                    itA = a.inter[mat1 * nm + mat2];
                    itB = a.inter[mat2 * nm + mat1];
                }
                if (FAST && !bnd && (FLUID || !((mk1 | mk2) & kMkElastic))) {
                    // both sides at once (one material, no elastic particle)
                    const float irj = frcp(xj.w);
                    PairSym o;
                    pair_vjp_sym(M1.mass, m2, rho1, xj.w, vi.w, vj.w, dq1, M1.visc + f2, M1.bulk_visc + bf2, itA,
                                 xj.x - xi.x, xj.y - xi.y, xj.z - xi.z, vj.x - vi.x, vj.y - vi.y, vj.z - vi.z,
                                 a1x, a1y, a1z, ws * in.wvx * irj, ws * in.wvy * irj, ws * in.wvz * irj, wp.w,
                                 ws * in.wd, hh, kn, a.phys.eps_hh, o);
                    for (int c = 0; c < 3; c++) {
                        G[c] -= o.gr[c];
                        Gv[c] -= o.gv[c];
                        F[c] += o.F[c];
                    }
                    Gr += o.gri;
                    pb[0] += o.dbulk;
                    pb[0] += o.dbulk;
                    pv[0] += o.dvisc;
                    pv[0] += o.dvisc;
                    return;
                }
                PairSide P;
                P.mi = M1.mass;
                P.mj = m2;
                P.rhoi = rho1;
                P.rhoj = xj.w;
                P.qi = vi.w;
                P.qj = vj.w;
                P.dqi = dq1;
                P.dqj = 0.0f;
                P.fi = M1.visc;
                P.fj = f2;
                P.bfi = M1.bulk_visc;
                P.bfj = bf2;
                P.eli = (mk1 & kMkElastic) != 0;
                P.elj = (mk2 & kMkElastic) != 0;
                P.bnd = bnd;
                P.same = mat1 == mat2;
                P.mati = mat1;
                P.matj = mat2;
                PairVjp o;
                // side a: this particle's own terms
                pair_vjp(P, itA, xj.x - xi.x, xj.y - xi.y, xj.z - xi.z, vj.x - vi.x,
                         vj.y - vi.y, vj.z - vi.z, a1x, a1y, a1z, wp.w, hh, kn, a.phys.eps_hh, o);
                for (int c = 0; c < 3; c++) {
                    G[c] -= o.gr[c];
                    Gv[c] -= o.gv[c];
                    F[c] += o.F[c];
                }
                Gr += o.gri;
                if constexpr (FAST) {
                    // mobile pairs are of one material; boundary pairs carry no
                    // viscosity terms
                    if (!bnd) {
                        pb[0] += o.dbulk;
                        pb[0] += o.dbulk;
                        pv[0] += o.dvisc;
                        pv[0] += o.dvisc;
                    }
                } else if (nm <= (uint32_t)kMatLds) {
                    pb[mat1] += o.dbulk;
                    pb[mat2] += o.dbulk;
                    pv[mat1] += o.dvisc;
                    pv[mat2] += o.dvisc;
                }
                if (!FLUID && !bnd) {
                    // side b: the neighbour's terms as seen from this particle
                    float4 wpj = make_float4(0.f, 0.f, 0.f, ws * in.wd);
                    const float4 wvj = make_float4(ws * in.wvx, ws * in.wvy, ws * in.wvz, 0.0f);
                    const float irj = frcp(xj.w);
                    PairSide Q;
                    Q.mi = m2;
                    Q.mj = M1.mass;
                    Q.rhoi = xj.w;
                    Q.rhoj = rho1;
                    Q.qi = vj.w;
                    Q.qj = vi.w;
                    Q.dqi = 0.0f;
                    Q.dqj = dq1;
                    Q.fi = f2;
                    Q.fj = M1.visc;
                    Q.bfi = bf2;
                    Q.bfj = M1.bulk_visc;
                    Q.eli = P.elj;
                    Q.elj = P.eli;
                    Q.bnd = false;
                    Q.same = P.same;
                    Q.mati = mat2;
                    Q.matj = mat1;
                    pair_vjp(Q, itB, xi.x - xj.x, xi.y - xj.y, xi.z - xj.z, vi.x - vj.x,
                             vi.y - vj.y, vi.z - vj.z, wvj.x * irj, wvj.y * irj, wvj.z * irj, wpj.w, hh, kn,
                             a.phys.eps_hh, o);
                    for (int c = 0; c < 3; c++) {
                        G[c] += o.gr[c];
                        Gv[c] += o.gv[c];
                    }
                    Gr += o.grj;
                }
            };
            if constexpr (REC) {
                // the forward stage's record of this particle's neighbours, streamed from
                // the record (rows of n entries: coalesced) in the recorded order; an entry
                // is loaded a pair ahead of its data, and an index past the count reads
                // the first row (in bounds; the host keeps the record below 4 GB)
                const char* nb = reinterpret_cast<const char*>(a.nbList);
                const uint32_t rcnt = a.nbCount[i], n32 = a.n;
                auto ent = [&](uint32_t k) {
                    return *reinterpret_cast<const uint32_t*>(nb + ((k < rcnt ? k : 0u) * n32 + i) * 4u);
                };
                if (rcnt) {
                    uint32_t ea = ent(0), eb = ent(1);
                    AdjIn pa = ld(ea), pbn;
                    for (uint32_t k = 0; k < rcnt; k += 2) {
                        const uint32_t ec = ent(k + 2);
                        pbn = ld(eb);
                        proc(ea, pa);
                        if (k + 1 >= rcnt) break;
                        const uint32_t ed = ent(k + 3);
                        pa = ld(ec);
                        proc(eb, pbn);
                        ea = ec;
                        eb = ed;
                    }
                }
            } else {
                uint32_t cx, cy, cz;
                cell_of(a.grid, xi.x, xi.y, xi.z, cx, cy, cz);
                const uint32_t sx = a.grid.sub[0], sy = a.grid.sub[1], sz = a.grid.sub[2];
                const uint32_t stampv = *a.stamp;
                const uint32_t xl = cx > 0 ? cx - 1 : 0, xh = min(cx + 1, sx - 1);
                // the candidate sweep of the forward kernel: branch-free batches, one drain site
                auto row_meta = [&](int rw, uint4& m) {
                    m = make_uint4(0, 0, 0, 0);
                    const int nz = (int)cz + rw / 3 - 1, ny = (int)cy + rw % 3 - 1;
                    if (rw < 9 && nz >= 0 && nz < (int)sz && ny >= 0 && ny < (int)sy) {
                        const uint32_t row = (uint32_t)ny + sy * (uint32_t)nz;
                        const uint32_t* ro = a.rowOff + (size_t)row * a.grid.rowstride;
                        const size_t cb = (size_t)row * sx;
                        const uint32_t st = a.rowStamp[row];
                        const uint32_t lo = ro[xl], hi = ro[xh + 1];
                        m.z = a.bP[cb + xl];
                        m.w = a.bP[cb + xh + 1];
                        if (st == stampv) {
                            m.x = lo;
                            m.y = hi;
                        }
                    }
                };
                uint32_t nl = 0;
                int rw = 0, part = 0;
                uint4 cur, nxtm;
                row_meta(0, cur);
                row_meta(1, nxtm);
                uint32_t q = cur.x, qend = cur.y;
                for (;;) {
                    while (nl + ARP_BATCH <= (uint32_t)kList) {
                        if (q >= qend) {
                            if (part == 0) {
                                part = 1;
                                q = cur.z;
                                qend = cur.w;
                                continue;
                            }
                            if (++rw >= 9) break;
                            cur = nxtm;
                            row_meta(rw + 1, nxtm);
                            part = 0;
                            q = cur.x;
                            qend = cur.y;
                            continue;
                        }
                        const bool bseg = part != 0;
                        const float4* base = bseg ? B : X;
                        const uint32_t tag = bseg ? kBnd : 0u;
                        uint32_t e[ARP_BATCH], ok[ARP_BATCH];
    #pragma unroll
                        for (int u = 0; u < ARP_BATCH; u++) {
                            const uint32_t qq = q + u;  // < qend + ARP_BATCH: inside the padding
                            const uint32_t inr = qq < qend ? 1u : 0u;
                            uint32_t v = qq, okv = inr;
                            if (!a.identity) {
                                const uint32_t pe = a.perm[bseg ? 0u : qq];
                                okv = inr & ((bseg || !(pe & kFlag)) ? 1u : 0u);
                                v = bseg ? qq : (inr ? (pe & ~kFlag) : 0u);
                            } else {
                                const uint32_t fl = a.sflag[bseg ? 0u : qq];
                                okv = inr & ((bseg || fl == 0u) ? 1u : 0u);
                            }
                            e[u] = v;
                            ok[u] = okv;
                        }
                        float4 xj[ARP_BATCH];
    #pragma unroll
                        for (int u = 0; u < ARP_BATCH; u++) xj[u] = base[e[u]];
    #pragma unroll
                        for (int u = 0; u < ARP_BATCH; u++) {
                            const float rx = xj[u].x - xi.x, ry = xj[u].y - xi.y, rz = xj[u].z - xi.z;
                            const uint32_t v = e[u] | tag;
                            const uint32_t in = ok[u] & (v != i ? 1u : 0u) &
                                                (hh - (rx * rx + ry * ry + rz * rz) > 0.0f ? 1u : 0u);
                            list[nl * 256 + tid] = v;
                            nl += in;
                        }
                        q += ARP_BATCH;
                    }
                    {
                        // (the next loads are unconditional: a lane at its last pair reloads its
                        // first entry rather than branching around the load)
                        AdjIn pa, pbn;
                        uint32_t ea = 0, eb = 0;
                        if (nl) {
                            ea = list[tid];
                            pa = ld(ea);
                        }
                        for (uint32_t k = 0; k < nl; k += 2) {
                            eb = list[(k + 1 < nl ? k + 1 : 0u) * 256 + tid];
                            pbn = ld(eb);
                            proc(ea, pa);
                            if (k + 1 >= nl) break;
                            ea = list[(k + 2 < nl ? k + 2 : 0u) * 256 + tid];
                            pa = ld(ea);
                            proc(eb, pbn);
                        }
                    }
                    nl = 0;
                    if (rw >= 9) break;
                }
            }
        }
        // pos' = v; vel' = F/rho + g
        Gv[0] += wp.x;
        Gv[1] += wp.y;
        Gv[2] += wp.z;
        Gr += -(wv.x * F[0] + wv.y * F[1] + wv.z * F[2]) * ir1 * ir1;
        const float4 gp = make_float4(G[0], G[1], G[2], Gr), gv = make_float4(Gv[0], Gv[1], Gv[2], 0.0f);
        if (a.Gp) {  // (kept when a later reverse stage reads it)
            reinterpret_cast<float4*>(a.Gp)[i] = gp;
            reinterpret_cast<float4*>(a.Gv)[i] = gv;
        }
        float4* LP = reinterpret_cast<float4*>(a.LamP) + i;
        float4* LV = reinterpret_cast<float4*>(a.LamV) + i;
        float4 lp, lv;
        if (a.LamInP) {
            lp = reinterpret_cast<const float4*>(a.LamInP)[i];
            lv = reinterpret_cast<const float4*>(a.LamInV)[i];
            if (a.WnP) {
                const float wl = a.wl, wg = a.wg;
                reinterpret_cast<float4*>(a.WnP)[i] =
                    make_float4(wl * lp.x + wg * gp.x, wl * lp.y + wg * gp.y, wl * lp.z + wg * gp.z, wl * lp.w + wg * gp.w);
                reinterpret_cast<float4*>(a.WnV)[i] =
                    make_float4(wl * lv.x + wg * gv.x, wl * lv.y + wg * gv.y, wl * lv.z + wg * gv.z, 0.0f);
            }
        } else {
            lp = *LP;
            lv = *LV;
        }
        lp.x += gp.x; lp.y += gp.y; lp.z += gp.z; lp.w += gp.w;
        lv.x += gv.x; lv.y += gv.y; lv.z += gv.z;
        if (a.permOut) {
            const uint32_t d = a.permOut[i];
            reinterpret_cast<float4*>(a.LamOutP)[d] = lp;
            reinterpret_cast<float4*>(a.LamOutV)[d] = lv;
        } else {
            *LP = lp;
            *LV = lv;
        }
    }
    // parameter partials: block sum in LDS, one row of partials per block
    if (nm <= (uint32_t)kMatLds) {
        for (uint32_t m = 0; m < nm; m++) {
            if (FAST && m != a.fastMat) continue;  // (uniform) the other materials' partials are 0
            float b = pb[FAST ? 0 : m], v = pv[FAST ? 0 : m];
            for (int o = 32; o >= 1; o >>= 1) {
                b += __shfl_xor(b, o);
                v += __shfl_xor(v, o);
            }
            if (__lane_id() == 0) {
                spg[tid >> 6][2 * m] = b;
                spg[tid >> 6][2 * m + 1] = v;
            }
        }
        __syncthreads();
        for (uint32_t m = tid; m < 2 * nm; m += 256)
            a.pgrad[(size_t)blockIdx.x * 2 * nm + m] += ((spg[0][m] + spg[1][m]) + spg[2][m]) + spg[3][m];
    }
}

// reverse stage: searching the grid / on the recorded neighbours; _g: general material
// tables (more than kMatLds materials or more than one mobile material)
extern "C" __global__ void __launch_bounds__(256, 3) k_fluid_adj(AdjArgs a) { fluid_adj_body<false, true>(a); }
extern "C" __global__ void __launch_bounds__(256, 3) k_fluid_adj_rec(AdjArgs a) { fluid_adj_body<true, true>(a); }
extern "C" __global__ void __launch_bounds__(256, 3) k_fluid_adj_g(AdjArgs a) { fluid_adj_body<false, false>(a); }
extern "C" __global__ void __launch_bounds__(256, 3) k_fluid_adj_rec_g(AdjArgs a) { fluid_adj_body<true, false>(a); }
// _f: scenes without elastic particles
extern "C" __global__ void __launch_bounds__(256, 3) k_fluid_adj_f(AdjArgs a) { fluid_adj_body<false, true, true>(a); }
extern "C" __global__ void __launch_bounds__(256, 4) k_fluid_adj_rec_f(AdjArgs a) { fluid_adj_body<true, true, true>(a); }

extern "C" __global__ void __launch_bounds__(256) k_adj_w(AdjWArgs a) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= a.n) return;
    float4 p = make_float4(0.f, 0.f, 0.f, 0.f), v = p;
    if (a.cl != 0.0f) {
        const float4 lp = reinterpret_cast<const float4*>(a.LP)[i], lv = reinterpret_cast<const float4*>(a.LV)[i];
        p = make_float4(a.cl * lp.x, a.cl * lp.y, a.cl * lp.z, a.cl * lp.w);
        v = make_float4(a.cl * lv.x, a.cl * lv.y, a.cl * lv.z, 0.f);
    }
    for (int k = 0; k < a.ng; k++) {
        const float4 gp = reinterpret_cast<const float4*>(a.GP[k])[i], gv = reinterpret_cast<const float4*>(a.GV[k])[i];
        const float c = a.cg[k];
        p.x += c * gp.x; p.y += c * gp.y; p.z += c * gp.z; p.w += c * gp.w;
        v.x += c * gv.x; v.y += c * gv.y; v.z += c * gv.z;
    }
    reinterpret_cast<float4*>(a.WP)[i] = p;
    reinterpret_cast<float4*>(a.WV)[i] = v;
}

// Velocity Verlet before the step's force evaluation: the kick base Z (in place) and the
// force point X from y and the last evaluation (VerletArgs).
extern "C" __global__ void __launch_bounds__(256) k_verlet_pre(VerletArgs a) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    const float dt = a.dt, h = 0.5f * dt;
    if (i < a.n) {
        float4 p = reinterpret_cast<const float4*>(a.posd)[i];
        float4 v = reinterpret_cast<const float4*>(a.velq)[i];
        const float4 dp = reinterpret_cast<const float4*>(a.Dp)[i];
        const float4 dv = reinterpret_cast<const float4*>(a.Dv)[i];
        const DMat& M = a.mats[a.mk[i] & kMatMask];
        v.x = v.x + h * dv.x;
        v.y = v.y + h * dv.y;
        v.z = v.z + h * dv.z;
        p.x = p.x + dt * v.x;
        p.y = p.y + dt * v.y;
        p.z = p.z + dt * v.z;
        p.w = p.w + h * dp.w;
        v.w = pressure_q(M, p.w);
        reinterpret_cast<float4*>(a.posd)[i] = p;
        reinterpret_cast<float4*>(a.velq)[i] = v;
        v.x = v.x + h * dv.x;
        v.y = v.y + h * dv.y;
        v.z = v.z + h * dv.z;
        p.w = p.w + h * dp.w;
        v.w = pressure_q(M, p.w);
        reinterpret_cast<float4*>(a.Xposd)[i] = p;
        reinterpret_cast<float4*>(a.Xvelq)[i] = v;
    }
    if (a.eps && i < a.ns)
        for (int q = 0; q < 16; q++) {
            const size_t k = 16 * (size_t)i + q;
            const float z = a.eps[k] + h * a.De[k];
            a.eps[k] = z;
            a.Xeps[k] = z + h * a.De[k];
        }
}

extern "C" __global__ void __launch_bounds__(256) k_adj_perm(AdjPermArgs a) {
    const uint32_t t = blockIdx.x * 256u + threadIdx.x;
    if (t >= a.n) return;
    const uint32_t d = a.src ? a.src[t] : t;
    reinterpret_cast<float4*>(a.outP)[d] = reinterpret_cast<const float4*>(a.inP)[t];
    reinterpret_cast<float4*>(a.outV)[d] = reinterpret_cast<const float4*>(a.inV)[t];
}

extern "C" __global__ void __launch_bounds__(256) k_grad_aos(GradAosArgs a) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= a.n) return;
    AosParticle* A = reinterpret_cast<AosParticle*>(a.aos) + a.pid[i];
    float4* P = reinterpret_cast<float4*>(a.P) + i;
    float4* V = reinterpret_cast<float4*>(a.V) + i;
    if (a.to_aos) {
        const float4 p = *P, v = *V;
        AosParticle o;
        o.den = p.w;
        o.mat = 0;
        o.solid_id = 0;
        o.pad = 0;
        o.pos = make_float4(p.x, p.y, p.z, 0.f);
        o.vel = make_float4(v.x, v.y, v.z, 0.f);
        *A = o;
    } else {
        const AosParticle o = *A;
        *P = make_float4(o.pos.x, o.pos.y, o.pos.z, o.den);
        *V = make_float4(o.vel.x, o.vel.y, o.vel.z, 0.f);
    }
}

extern "C" __global__ void __launch_bounds__(256) k_reduce(ReduceArgs a) {
    // one workgroup per output entry; fixed summation order
    __shared__ float s[256];
    const uint32_t w = blockIdx.x, tid = threadIdx.x;
    float acc = 0.0f;
    for (uint32_t b = tid; b < a.nblocks; b += 256) acc += a.part[(size_t)b * a.width + w];
    s[tid] = acc;
    __syncthreads();
    for (uint32_t o = 128; o > 0; o >>= 1) {
        if (tid < o) s[tid] += s[tid + o];
        __syncthreads();
    }
    if (tid == 0) a.out[w] += s[0];
}

// ------------------------------------------------------------------------------------
// adjoint of the solid part (reverse mode of strain.rs:139-160, solid.rs:105-168, 334-450 and
// the plastic flow of solid.rs:499-518)
// ------------------------------------------------------------------------------------
extern "C" __global__ void __launch_bounds__(256) k_inv(InvArgs a) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i < a.n) a.inv[a.pid[i]] = i;
}

// d(a . f)/dr for the self-contact force f(r) = cs(l) r / l, cs = -5000 sqrt(max(rc - l, 0) rg)
__device__ __forceinline__ float3 contact_vjp(float ax, float ay, float az, float rx, float ry, float rz, float rc,
                                             float rg) {
    const float l = sqrtf(rx * rx + ry * ry + rz * rz);
    const float dr = fmaxf(rc - l, 0.0f);
    const float sq = sqrtf(dr * rg);
    const float cs = -5000.0f * sq;
    const float dcs = sq > 0.0f ? 2500.0f * rg / sq : 0.0f;  // d cs / d l
    const float ar = ax * rx + ay * ry + az * rz;
    const float k1 = cs / l, k2 = ar * (dcs / l - cs / (l * l)) / l;
    return make_float3(k1 * ax + k2 * rx, k1 * ay + k2 * ry, k1 * az + k2 * rz);
}

// Hourglass force of target 1 on the pair (1, 2) (solid.rs:432-443):
//   f = -sc dx/(l + eps h),  sc = K T Q W,  K = 1e8 m1 m2/(d1 d2),
//   T = (err . dx)/2/(l + 10 eps h), Q = 1/(l^2 + eps h^2), err = ((F1+F2)/2) dX - dx.
// Given the cotangent a of f: returns f, and adds d(a.f)/d dx to gdx, d/d(fa rows) (the
// cotangent of (F1+F2)/2) to gfa, and d/dd1, d/dd2.
struct HgOut {
    float f[3], gdx[3], gfa[9], gd1, gd2;
};
__device__ __forceinline__ void hourglass_vjp(const float3* F1, const float3* F2, const float dX[3], const float dx[3],
                                              float m1, float m2, float d1, float d2, float W, float h, float ax,
                                              float ay, float az, HgOut& o) {
    const float EPS = 0.1f;  // EPSILON of the reference shaders (fluid.rs:9, solid.rs:9)
    float fa[9];
    for (int r = 0; r < 3; r++) {
        const float3 x = F1[r], y = F2[r];
        fa[3 * r + 0] = 0.5f * (x.x + y.x);
        fa[3 * r + 1] = 0.5f * (x.y + y.y);
        fa[3 * r + 2] = 0.5f * (x.z + y.z);
    }
    float err[3];
    for (int r = 0; r < 3; r++) err[r] = fa[3 * r] * dX[0] + fa[3 * r + 1] * dX[1] + fa[3 * r + 2] * dX[2] - dx[r];
    const float l = sqrtf(dx[0] * dx[0] + dx[1] * dx[1] + dx[2] * dx[2]);
    const float ed = err[0] * dx[0] + err[1] * dx[1] + err[2] * dx[2];
    const float K = 100000000.0f * m1 * m2 / (d1 * d2);
    const float l10 = l + 10.0f * EPS * h, l1 = l + EPS * h, qd = l * l + EPS * h * h;
    const float T = 0.5f * ed / l10, Q = 1.0f / qd;
    const float sc = K * T * Q * W;
    const float a[3] = {ax, ay, az};
    const float adx = ax * dx[0] + ay * dx[1] + az * dx[2];
    const float au = adx / l1;  // a . u, u = dx / l1
    for (int r = 0; r < 3; r++) o.f[r] = -sc * dx[r] / l1;
    // d(a.f) = -(a.u) d sc - sc a.du
    const float cT = -au * K * Q * W, cQ = -au * K * T * W;
    const float il = l > 0.0f ? 1.0f / l : 0.0f;
    for (int r = 0; r < 3; r++) {
        const float dT = 0.5f * (err[r] - dx[r]) / l10 - 0.5f * ed / (l10 * l10) * dx[r] * il;
        const float dQ = -Q * Q * 2.0f * dx[r];
        const float du = a[r] / l1 - adx * dx[r] * il / (l1 * l1);
        o.gdx[r] += cT * dT + cQ * dQ - sc * du;
    }
    // fa enters through err: d ed / d fa[r][k] = dX[k] dx[r]
    for (int r = 0; r < 3; r++)
        for (int k = 0; k < 3; k++) o.gfa[3 * r + k] += cT * 0.5f / l10 * dX[k] * dx[r];
    o.gd1 += au * sc / d1;
    o.gd2 += au * sc / d2;
}

__device__ __forceinline__ float4 ld_w(const float* base, uint32_t i, float ws) {
    const float4 v = reinterpret_cast<const float4*>(base)[i];
    return make_float4(ws * v.x, ws * v.y, ws * v.z, ws * v.w);
}

extern "C" __global__ void __launch_bounds__(256) k_solid_adj1(SolidAdjArgs aa) {
    const SolidArgs& a = aa.f;
    const uint32_t s = blockIdx.x * 256u + threadIdx.x;
    if (s >= a.ns) return;
    float* fb = aa.Fbar + kFbarStride * (size_t)s;
    const uint32_t p = a.part[s];
    const uint32_t i = p == kInvalid ? 0u : a.inv[p];
    if (p == kInvalid || !(a.mk[i] & kMkElastic)) {
        for (int q = 0; q < 18; q++) fb[q] = 0.0f;
        return;
    }
    const uint32_t mk = a.mk[i], mat = mk & kMatMask;
    const DMat& M = a.mats[mat];
    const float4* X = reinterpret_cast<const float4*>(a.Xposd);
    const float4* XV = reinterpret_cast<const float4*>(a.Xvelq);
    const float4* REF = reinterpret_cast<const float4*>(a.ref);
    const float h = a.phys.h;
    const uint32_t dim = a.grid.dim_n;
    const float ws = aa.wscale;
    const float4 xi = X[i];
    const float m1 = M.mass, d1 = xi.w, V1 = m1 / d1;
    const float4 wpi = ld_w(aa.Wp, i, ws), wvi = ld_w(aa.Wv, i, ws);
    const float ai[3] = {wvi.x / d1, wvi.y / d1, wvi.z / d1};
    const float zm = dim > 2 ? 1.0f : 0.0f;  // the elastic force drops z below 3-D
    const float3* A1 = reinterpret_cast<const float3*>(a.A) + 3 * (size_t)s;
    const float3* F1 = reinterpret_cast<const float3*>(a.Fm) + 3 * (size_t)s;
    const float4 Xs = REF[s];
    const bool valid_i = a.sflag[i] == 0u;
    uint32_t ci[3] = {0, 0, 0};
    cell_of(a.grid, xi.x, xi.y, xi.z, ci[0], ci[1], ci[2]);
    float Ab[9] = {0}, Fhb[9] = {0}, xb[3] = {0, 0, 0}, rb = 0.0f, fs[3] = {0, 0, 0};
    const float4* EG = reinterpret_cast<const float4*>(a.eG);
    const float3* AA = reinterpret_cast<const float3*>(a.A);
    const float4* WVv = reinterpret_cast<const float4*>(aa.Wv);
    const uint32_t q0 = a.eOff[s], q1 = a.eOff[s + 1];
    // ARP_ADJ_GROUP entries at a time, each gather level issued for the group (as k_solid)
    for (uint32_t qb = q0; qb < q1; qb += ARP_ADJ_GROUP) {
        uint32_t sq[ARP_ADJ_GROUP], jq[ARP_ADJ_GROUP];
        float4 gq[ARP_ADJ_GROUP], xq[ARP_ADJ_GROUP], wq[ARP_ADJ_GROUP];
        float3 Aq[ARP_ADJ_GROUP][3];
#pragma unroll
        for (int k = 0; k < ARP_ADJ_GROUP; k++) {
            const uint32_t q = min(qb + (uint32_t)k, q1 - 1u);
            sq[k] = a.eSid[q];
            jq[k] = a.ePid[q];
            gq[k] = EG[q];
        }
#pragma unroll
        for (int k = 0; k < ARP_ADJ_GROUP; k++) {
            jq[k] = a.inv[jq[k]];
#pragma unroll
            for (int r = 0; r < 3; r++) Aq[k][r] = AA[3 * (size_t)sq[k] + r];
        }
#pragma unroll
        for (int k = 0; k < ARP_ADJ_GROUP; k++) {
            xq[k] = X[jq[k]];
            wq[k] = WVv[jq[k]];
        }
#pragma unroll
        for (int k = 0; k < ARP_ADJ_GROUP; k++) {
        if (qb + (uint32_t)k >= q1) break;
        const uint32_t s2 = sq[k], j = jq[k];
        const float4 g = gq[k], xj = xq[k];
        const float dj = xj.w, V2 = m1 / dj, c = -V1 * V2;
        const float3* A2 = Aq[k];
        const float gg[3] = {g.x, g.y, g.z};
        float e[3];
        for (int r = 0; r < 3; r++) {
            const float3 x = A1[r], y = A2[r];
            e[r] = (x.x + y.x) * gg[0] + (x.y + y.y) * gg[1] + (x.z + y.z) * gg[2];
        }
        fs[0] += c * e[0];
        fs[1] += c * e[1];
        fs[2] += zm * c * e[2];
        const float4 wvj = make_float4(ws * wq[k].x, ws * wq[k].y, ws * wq[k].z, ws * wq[k].w);
        const float aj[3] = {wvj.x / dj, wvj.y / dj, wvj.z / dj};
        const float aim[3] = {ai[0], ai[1], zm * ai[2]}, ajm[3] = {aj[0], aj[1], zm * aj[2]};
        // own force c (A1 + A2) g and the partner's c (A2 + A1)(-g): both hold A1
        for (int r = 0; r < 3; r++)
            for (int k = 0; k < 3; k++) Ab[3 * r + k] += c * (aim[r] - ajm[r]) * gg[k];
        const float ae_i = aim[0] * e[0] + aim[1] * e[1] + aim[2] * e[2];
        const float ae_j = ajm[0] * e[0] + ajm[1] * e[1] + ajm[2] * e[2];
        rb += (-c / d1) * ae_i + (-c / d1) * (-ae_j);
        // hourglass, own term and the partner's term on the same pair
        if (g.w != 0.0f && valid_i && a.sflag[j] == 0u) {
            uint32_t cj[3] = {0, 0, 0};
            cell_of(a.grid, xj.x, xj.y, xj.z, cj[0], cj[1], cj[2]);
            bool adj = true;
            for (int d = 0; d < 3; d++) adj &= (ci[d] + 1 >= cj[d]) && (cj[d] + 1 >= ci[d]);
            if (adj) {
                const float4 X2 = REF[s2];
                const float3* F2 = reinterpret_cast<const float3*>(a.Fm) + 3 * (size_t)s2;
                const float dX[3] = {X2.x - Xs.x, X2.y - Xs.y, X2.z - Xs.z};
                const float dx[3] = {xj.x - xi.x, xj.y - xi.y, xj.z - xi.z};
                HgOut o = {};
                hourglass_vjp(F1, F2, dX, dx, m1, m1, d1, dj, g.w, h, ai[0], ai[1], ai[2], o);
                for (int r = 0; r < 3; r++) {
                    fs[r] += o.f[r];
                    xb[r] -= o.gdx[r];
                }
                rb += o.gd1;
                for (int q2 = 0; q2 < 9; q2++) Fhb[q2] += 0.5f * o.gfa[q2];
                const float mdX[3] = {-dX[0], -dX[1], -dX[2]}, mdx[3] = {-dx[0], -dx[1], -dx[2]};
                HgOut o2 = {};
                hourglass_vjp(F2, F1, mdX, mdx, m1, m1, dj, d1, g.w, h, aj[0], aj[1], aj[2], o2);
                for (int r = 0; r < 3; r++) xb[r] += o2.gdx[r];
                rb += o2.gd2;
                for (int q2 = 0; q2 < 9; q2++) Fhb[q2] += 0.5f * o2.gfa[q2];
            }
        }
        }
    }
    // self contact (solid.rs:426-450): own force and the partners' forces on this particle
    if (!M.plastic && valid_i) {
        const float r_cut = 8.0f * h, r_geom = h * h / r_cut;
        const uint32_t sx = a.grid.sub[0], sy = a.grid.sub[1], sz = a.grid.sub[2];
        const uint32_t xl = ci[0] > 0 ? ci[0] - 1 : 0, xh = min(ci[0] + 1, sx - 1);
        const uint32_t stampv = *a.stamp;
        const float reach = contact_reach(a, xi, r_cut);
        const float u0 = xi.x - Xs.x, u1 = xi.y - Xs.y, u2 = xi.z - Xs.z;
        for (int dz = -1; dz <= 1; dz++) {
            const int nz = (int)ci[2] + dz;
            if (nz < 0 || nz >= (int)sz) continue;
            for (int dy = -1; dy <= 1; dy++) {
                const int ny = (int)ci[1] + dy;
                if (ny < 0 || ny >= (int)sy) continue;
                const uint32_t row = (uint32_t)ny + sy * (uint32_t)nz;
                if (a.rowStamp[row] != stampv) continue;
                if (contact_row_clear(a, row, u0, u1, u2, reach)) continue;
                const uint32_t* ro = a.rowOff + (size_t)row * a.grid.rowstride;
                for (uint32_t q = ro[xl]; q < ro[xh + 1]; q++) {
                    const uint32_t pe = a.perm[q];
                    if (pe & kFlag) continue;
                    const uint32_t mk2 = a.mk[pe];
                    if ((mk2 & kMatMask) != mat || !(mk2 & kMkSolid)) continue;
                    const uint32_t s2 = a.sidByPid[a.pid[pe]];
                    const float4 X2 = REF[s2];
                    const float dX0 = X2.x - Xs.x, dX1 = X2.y - Xs.y, dX2 = X2.z - Xs.z;
                    if (!(dX0 * dX0 + dX1 * dX1 + dX2 * dX2 > r_cut * r_cut)) continue;
                    const float4 xj = X[pe];
                    const float rx = xj.x - xi.x, ry = xj.y - xi.y, rz = xj.z - xi.z;
                    const float rr = rx * rx + ry * ry + rz * rz;
                    if (!(rr < r_cut * r_cut)) continue;
                    const float l = sqrtf(rr);
                    const float cs = -5000.0f * sqrtf(fmaxf(r_cut - l, 0.0f) * r_geom);
                    fs[0] += cs * rx / l;
                    fs[1] += cs * ry / l;
                    fs[2] += cs * rz / l;
                    const float3 go = contact_vjp(ai[0], ai[1], ai[2], rx, ry, rz, r_cut, r_geom);
                    const float4 wvj = ld_w(aa.Wv, pe, ws);
                    const float dj = xj.w;
                    const float3 gp = contact_vjp(wvj.x / dj, wvj.y / dj, wvj.z / dj, -rx, -ry, -rz, r_cut, r_geom);
                    xb[0] += gp.x - go.x;
                    xb[1] += gp.y - go.y;
                    xb[2] += gp.z - go.z;
                }
            }
        }
    }
    // vel' += fs / d1, pos' += v (dest2, shaders/mod.rs:79)
    rb += -(wvi.x * fs[0] + wvi.y * fs[1] + wvi.z * fs[2]) / (d1 * d1);
    float4* GP = reinterpret_cast<float4*>(aa.Gp) + i;
    float4* GV = reinterpret_cast<float4*>(aa.Gv) + i;
    float4* LP = reinterpret_cast<float4*>(aa.LamP) + i;
    float4* LV = reinterpret_cast<float4*>(aa.LamV) + i;
    float4 gp = *GP, gv = *GV, lp = *LP, lv = *LV;
    gp.x += xb[0]; gp.y += xb[1]; gp.z += xb[2]; gp.w += rb;
    gv.x += wpi.x; gv.y += wpi.y; gv.z += wpi.z;
    lp.x += xb[0]; lp.y += xb[1]; lp.z += xb[2]; lp.w += rb;
    lv.x += wpi.x; lv.y += wpi.y; lv.z += wpi.z;
    *GP = gp; *GV = gv; *LP = lp; *LV = lv;

    // strain adjoint: A = (F S^T) L^-1, S = hooke(E) + hooke(E'), E = (F^T F - I)/2 + eps,
    // E' = (F'^T F + F^T F')/2
    const float4* XVv = XV;
    const float4 vi = XVv[i];
    float F[9] = {0}, Fd[9] = {0};
    const float4* G = reinterpret_cast<const float4*>(a.sG);
    // F and F' of the recorded stage as its k_strain formed them (the same sums in the same
    // order, so the same values); without them, formed again from the strain neighbours
    const uint32_t p0 = a.sOff[s], p1 = a.Fd ? p0 : a.sOff[s + 1];
    if (a.Fd) {
        const float3* Fr = reinterpret_cast<const float3*>(a.Fm) + 3 * (size_t)s;
        const float3* Fdr = reinterpret_cast<const float3*>(a.Fd) + 3 * (size_t)s;
        for (int r = 0; r < 3; r++) {
            const float3 f = Fr[r], fd = Fdr[r];
            F[3 * r] = f.x, F[3 * r + 1] = f.y, F[3 * r + 2] = f.z;
            Fd[3 * r] = fd.x, Fd[3 * r + 1] = fd.y, Fd[3 * r + 2] = fd.z;
        }
    }
    for (uint32_t qb = p0; qb < p1; qb += ARP_ST_GROUP) {  // (grouped as in k_strain)
        uint32_t jj[ARP_ST_GROUP];
        float4 gq[ARP_ST_GROUP], xq[ARP_ST_GROUP], vq[ARP_ST_GROUP];
#pragma unroll
        for (int k = 0; k < ARP_ST_GROUP; k++) {
            const uint32_t q = min(qb + (uint32_t)k, p1 - 1u);
            jj[k] = a.sPid[q];
            gq[k] = G[q];
        }
#pragma unroll
        for (int k = 0; k < ARP_ST_GROUP; k++) jj[k] = a.inv[jj[k]];
#pragma unroll
        for (int k = 0; k < ARP_ST_GROUP; k++) {
            xq[k] = X[jj[k]];
            vq[k] = XVv[jj[k]];
        }
#pragma unroll
        for (int k = 0; k < ARP_ST_GROUP; k++) {
            if (qb + (uint32_t)k >= p1) break;
            const float4 g = gq[k], xj = xq[k], vj = vq[k];
            const float dx[3] = {xj.x - xi.x, xj.y - xi.y, xj.z - xi.z};
            const float dv[3] = {vj.x - vi.x, vj.y - vi.y, vj.z - vi.z};
            const float gg[3] = {g.x, g.y, g.z};
            for (int r = 0; r < 3; r++)
                for (int c = 0; c < 3; c++) {
                    F[3 * r + c] += g.w * (dx[r] * gg[c]);
                    Fd[3 * r + c] += g.w * (dv[r] * gg[c]);
                }
        }
    }
    M4 F4 = m4_zero(), Fd4 = m4_zero(), Ab4 = m4_zero(), Fh4 = m4_zero();
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) {
            F4.m[4 * c + r] = F[3 * r + c];
            Fd4.m[4 * c + r] = Fd[3 * r + c];
            Ab4.m[4 * c + r] = Ab[3 * r + c];
            Fh4.m[4 * c + r] = Fhb[3 * r + c];
        }
    const M4 eps = ld_m4(a.Xeps, s);
    const float ks = a.kscale ? a.kscale[s] : 1.0f;
    const float lam = ks * M.lambda, mu = ks * M.mu;
    const int order = M.strain_order;
    const M4 E = m4_add(m4_seth_hill(F4, order, dim), eps);
    const M4 Ed = m4_seth_hill_rate(F4, Fd4, order, dim);
    const M4 S = m4_add(m4_hooke(E, lam, mu, dim), m4_hooke(Ed, M.lambda_d, M.mu_d, dim));
    const M4 Li = ld_m4(a.Linv, s);
    const M4 Pb = m4_mul(Ab4, m4_trans(Li));
    M4 Sb = m4_mul(m4_trans(Pb), F4);
    M4 Fb = m4_add(m4_mul(Pb, S), Fh4);
    // plastic flow (solid.rs:499-518): rate = (|S| - A)/T S/|S| when |S|^2 >= A^2, with A the
    // yield stress, constant since eq_plastic_strain_sq carries the factor (dim-1)/dim == 0;
    // its cotangent is this stage's share of the plastic strain's cotangents
    const bool plast = M.plastic && aa.lamE;
    if (plast) {
        M4 rbar = m4_scale(aa.wcl, ld_m4(aa.lamE, s));
        for (int t = aa.stage + 1; t < aa.nstages; t++) rbar = m4_add(rbar, m4_scale(aa.wcg[t], ld_m4(aa.gE[t], s)));
        M4 dev = eps;
        const float tr = m4_trace(eps, dim);
        for (int c = 0; c < 4; c++) dev.m[5 * c] = eps.m[5 * c] - tr / (float)dim;
        const float eq = (float)((dim - 1) / dim) * m4_dot(dev, dev);
        const float ys = (M.yield + M.hard * powf(eq, M.hard_exp / 2.0f)) * (1.0f + M.kin_hard * logf(1.0f));
        const float sn2 = m4_dot(S, S);
        if (ys * ys <= sn2) {
            // d rate = c1 dS + c2 (S . dS) S
            const float n = sqrtf(sn2);
            const float c1 = (n - ys) / (M.relax * n), c2 = ys / (M.relax * n * n * n);
            const float sr = m4_dot(S, rbar);
            for (int q = 0; q < 16; q++) Sb.m[q] += c1 * rbar.m[q] + c2 * sr * S.m[q];
        }
    }
    const float trS = Sb.m[0] + Sb.m[5] + Sb.m[10] + Sb.m[15];
    M4 Eb, Edb;
    for (int c = 0; c < 4; c++)
        for (int r = 0; r < 4; r++) {
            const float dg = (r == c && (uint32_t)r < dim) ? trS : 0.0f;
            Eb.m[4 * c + r] = lam * dg + 2.0f * mu * Sb.m[4 * c + r];
            Edb.m[4 * c + r] = M.lambda_d * dg + 2.0f * M.mu_d * Sb.m[4 * c + r];
        }
    M4 Fdb = m4_zero();
    if (order == 2) {  // E = (F^T F - I)/2, E' = (F'^T F + F^T F')/2
        const M4 EbS = m4_add(Eb, m4_trans(Eb)), EdbS = m4_add(Edb, m4_trans(Edb));
        Fb = m4_add(Fb, m4_add(m4_scale(0.5f, m4_mul(F4, EbS)), m4_scale(0.5f, m4_mul(Fd4, EdbS))));
        Fdb = m4_scale(0.5f, m4_mul(F4, EdbS));
    } else if (aa.Ebuf) {
        // other orders: E, E' cotangents and F' for k_solid_adj_strain, which adds the
        // Seth-Hill terms to F-bar and forms F'-bar
        float4* eb = reinterpret_cast<float4*>(aa.Ebuf) + 12 * (size_t)s;
        for (int q = 0; q < 4; q++) {
            eb[q] = make_float4(Eb.m[4 * q], Eb.m[4 * q + 1], Eb.m[4 * q + 2], Eb.m[4 * q + 3]);
            eb[4 + q] = make_float4(Edb.m[4 * q], Edb.m[4 * q + 1], Edb.m[4 * q + 2], Edb.m[4 * q + 3]);
        }
        eb[8] = make_float4(Fd[0], Fd[1], Fd[2], Fd[3]);
        eb[9] = make_float4(Fd[4], Fd[5], Fd[6], Fd[7]);
        eb[10] = make_float4(Fd[8], 0.f, 0.f, 0.f);
    }
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) {
            fb[3 * r + c] = Fb.m[4 * c + r];
            fb[9 + 3 * r + c] = Fdb.m[4 * c + r];
        }
    float* sp = aa.spar + 4 * (size_t)s;
    sp[0] += m4_trace(E, dim) * trS;
    sp[1] += 2.0f * m4_dot(Sb, E);
    sp[2] += m4_trace(Ed, dim) * trS;
    sp[3] += 2.0f * m4_dot(Sb, Ed);
    if (plast) {
        // the stage input's plastic strain enters E (E = seth_hill(F) + eps): its cotangent
        // is E-bar; the last reverse stage folds every stage's into dL/d y_n
        float4* g = reinterpret_cast<float4*>(aa.gE[aa.stage]) + 4 * (size_t)s;
        for (int q = 0; q < 4; q++) g[q] = make_float4(Eb.m[4 * q], Eb.m[4 * q + 1], Eb.m[4 * q + 2], Eb.m[4 * q + 3]);
        if (aa.stage == 0) {
            M4 lam = m4_add(ld_m4(aa.lamE, s), Eb);
            for (int t = 1; t < aa.nstages; t++) lam = m4_add(lam, ld_m4(aa.gE[t], s));
            float4* l = reinterpret_cast<float4*>(aa.lamE) + 4 * (size_t)s;
            for (int q = 0; q < 4; q++) l[q] = make_float4(lam.m[4 * q], lam.m[4 * q + 1], lam.m[4 * q + 2], lam.m[4 * q + 3]);
        }
    }
}

// ---- forward-mode tangents of the Seth-Hill strain and its rate (strain orders != 2) ----
// A matrix and its tangent; each operation carries the derivative of the forward one
// (d(AB) = dA B + A dB, d(A^-1) = -A^-1 dA A^-1), so that the functions below give the exact
// directional derivative of m4_seth_hill / m4_seth_hill_rate as they compute.
struct DM4 {
    M4 v, d;
};
__device__ __forceinline__ DM4 dm4_mul(const DM4& A, const DM4& B) {
    DM4 r;
    r.v = m4_mul(A.v, B.v);
    r.d = m4_add(m4_mul(A.d, B.v), m4_mul(A.v, B.d));
    return r;
}
__device__ __forceinline__ DM4 dm4_add(const DM4& A, const DM4& B) {
    return DM4{m4_add(A.v, B.v), m4_add(A.d, B.d)};
}
__device__ __forceinline__ DM4 dm4_scale(float c, const DM4& A) { return DM4{m4_scale(c, A.v), m4_scale(c, A.d)}; }
__device__ __forceinline__ DM4 dm4_trans(const DM4& A) { return DM4{m4_trans(A.v), m4_trans(A.d)}; }
__device__ DM4 dm4_inv(const DM4& A) {
    DM4 r;
    r.v = m4_inv(A.v);
    r.d = m4_scale(-1.0f, m4_mul(m4_mul(r.v, A.d), r.v));
    return r;
}
__device__ DM4 dm4_cauchy(const DM4& F, uint32_t dim) {
    DM4 C = dm4_mul(dm4_trans(F), F);
#pragma unroll
    for (uint32_t i = 0; i < 4; i++)
        if (i >= dim) {
            C.v.m[5 * i] = 1.0f;
            C.d.m[5 * i] = 0.0f;
        }
    return C;
}
// the tangent of m4_seth_hill (solid.rs:105-168)
__device__ M4 dm4_seth_hill(const DM4& F, int two_m, uint32_t dim) {
    const DM4 C = dm4_cauchy(F, dim);
    const DM4 I{m4_ident(), m4_zero()};
    if (two_m == 0) {
        DM4 sum{m4_zero(), m4_zero()}, term = dm4_scale(-1.0f, I), AmI;
        for (int q = 0; q < 16; q++) {
            AmI.v.m[q] = -1.0f * (C.v.m[q] - I.v.m[q]);
            AmI.d.m[q] = -C.d.m[q];
        }
        for (int i = 1; i <= 5; i++) {
            term = dm4_mul(term, AmI);
            for (int q = 0; q < 16; q++) {
                sum.v.m[q] += term.v.m[q] / (float)i;
                sum.d.m[q] += term.d.m[q] / (float)i;
            }
        }
        return m4_scale(0.5f, sum.d);
    }
    const uint32_t m = (uint32_t)abs(two_m) >> 1;
    DM4 Cm = I;
    for (uint32_t i = 0; i < m; i++) Cm = dm4_mul(Cm, C);
    if ((abs(two_m) & 1) == 1) {
        DM4 Y = C, Z = I;
        for (int i = 0; i < 15; i++) {
            const DM4 iZ = dm4_inv(Z), iY = dm4_inv(Y);
            const DM4 Yn = dm4_scale(0.5f, dm4_add(Y, iZ)), Zn = dm4_scale(0.5f, dm4_add(Z, iY));
            Y = Yn;
            Z = Zn;
        }
        Cm = dm4_mul(Cm, Y);
    }
    if (two_m < 0) Cm = dm4_inv(Cm);
    return m4_scale(1.0f / (float)two_m, Cm.d);
}
// the tangent of m4_seth_hill_rate
__device__ M4 dm4_seth_hill_rate(const DM4& F, const DM4& Fd, int two_m, uint32_t dim) {
    DM4 dC = dm4_mul(dm4_trans(Fd), F);
    dC = dm4_add(dC, dm4_trans(dC));
    const DM4 C = dm4_cauchy(F, dim);
    if (two_m == 0) return dm4_mul(dm4_scale(0.5f, dC), dm4_inv(C)).d;
    const uint32_t m = (uint32_t)abs(two_m) >> 1;
    DM4 sr{m4_zero(), m4_zero()};
    for (uint32_t i = 0; i < m; i++) {
        DM4 term{m4_ident(), m4_zero()};
        for (uint32_t j = 0; j < m; j++) {
            if (j == i)
                term = dm4_mul(term, dC);
            else
                term = dm4_mul(term, C);
        }
        sr = dm4_add(sr, term);
    }
    sr = dm4_scale(1.0f / (float)two_m, sr);
    if (two_m < 0) {
        const DM4 iC = dm4_inv(C);
        return dm4_mul(dm4_mul(dm4_scale(-1.0f, iC), sr), iC).d;
    }
    return sr.d;
}

// Completes F-bar and forms F'-bar for solids of strain_order != 2 from the E, E' cotangents
// k_solid_adj1 left: F-bar += (dE/dF)^T E-bar + (dE'/dF)^T E'-bar, F'-bar = (dE'/dF')^T E'-bar,
// one forward-mode tangent per entry of F (and of F').
extern "C" __global__ void __launch_bounds__(256) k_solid_adj_strain(SolidAdjArgs aa) {
    const SolidArgs& a = aa.f;
    const uint32_t s = blockIdx.x * 256u + threadIdx.x;
    if (s >= a.ns || !aa.Ebuf) return;
    const uint32_t p = a.part[s];
    if (p == kInvalid) return;
    const uint32_t i = a.inv[p];
    if (!(a.mk[i] & kMkElastic)) return;
    const DMat& M = a.mats[a.mk[i] & kMatMask];
    const int order = M.strain_order;
    if (order == 2) return;
    const uint32_t dim = a.grid.dim_n;
    const float4* eb = reinterpret_cast<const float4*>(aa.Ebuf) + 12 * (size_t)s;
    M4 Eb, Edb;
    for (int q = 0; q < 4; q++) {
        const float4 x = eb[q], y = eb[4 + q];
        Eb.m[4 * q] = x.x, Eb.m[4 * q + 1] = x.y, Eb.m[4 * q + 2] = x.z, Eb.m[4 * q + 3] = x.w;
        Edb.m[4 * q] = y.x, Edb.m[4 * q + 1] = y.y, Edb.m[4 * q + 2] = y.z, Edb.m[4 * q + 3] = y.w;
    }
    const float fd9[9] = {eb[8].x, eb[8].y, eb[8].z, eb[8].w, eb[9].x, eb[9].y, eb[9].z, eb[9].w, eb[10].x};
    const float3* F1 = reinterpret_cast<const float3*>(a.Fm) + 3 * (size_t)s;
    M4 F4 = m4_zero(), Fd4 = m4_zero();
    for (int r = 0; r < 3; r++) {
        const float3 row = F1[r];
        F4.m[r] = row.x, F4.m[4 + r] = row.y, F4.m[8 + r] = row.z;
        for (int c = 0; c < 3; c++) Fd4.m[4 * c + r] = fd9[3 * r + c];
    }
    float* fb = aa.Fbar + kFbarStride * (size_t)s;
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) {
            DM4 F{F4, m4_zero()}, Fd{Fd4, m4_zero()};
            F.d.m[4 * c + r] = 1.0f;
            const M4 dE = dm4_seth_hill(F, order, dim);
            const M4 dEd = dm4_seth_hill_rate(F, Fd, order, dim);
            fb[3 * r + c] += m4_dot(Eb, dE) + m4_dot(Edb, dEd);
            DM4 G{F4, m4_zero()}, Gd{Fd4, m4_zero()};
            Gd.d.m[4 * c + r] = 1.0f;
            fb[9 + 3 * r + c] = m4_dot(Edb, dm4_seth_hill_rate(G, Gd, order, dim));
        }
}

extern "C" __global__ void __launch_bounds__(256) k_solid_adj2(SolidAdjArgs aa) {
    const SolidArgs& a = aa.f;
    const uint32_t s = blockIdx.x * 256u + threadIdx.x;
    if (s >= a.ns) return;
    const uint32_t p = a.part[s];
    if (p == kInvalid) return;
    const uint32_t i = a.inv[p];
    if (!(a.mk[i] & kMkElastic)) return;
    const DMat& M = a.mats[a.mk[i] & kMatMask];
    const float Vp = M.mass / M.start_den;  // this particle's volume in its neighbours' F
    const float* fb = aa.Fbar + kFbarStride * (size_t)s;
    float xb[3] = {0, 0, 0}, vb[3] = {0, 0, 0};
    const float4* G2 = reinterpret_cast<const float4*>(a.sG2);
    const float4* G = reinterpret_cast<const float4*>(a.sG);
    float f1[18];  // this solid's F, F' cotangents
    {
        const float4* fb4 = reinterpret_cast<const float4*>(fb);
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const float4 v = fb4[q];
            f1[4 * q] = v.x, f1[4 * q + 1] = v.y, f1[4 * q + 2] = v.z, f1[4 * q + 3] = v.w;
        }
        const float2 v = reinterpret_cast<const float2*>(fb)[8];
        f1[16] = v.x, f1[17] = v.y;
    }
    const uint32_t q0 = a.sOff[s], q1 = a.sOff[s + 1];
    for (uint32_t qb = q0; qb < q1; qb += ARP_ADJ_GROUP) {
        uint32_t sq[ARP_ADJ_GROUP];
        float4 gq[ARP_ADJ_GROUP], g2q[ARP_ADJ_GROUP], Fq[ARP_ADJ_GROUP][4];
        float2 Ft[ARP_ADJ_GROUP];  // (floats 16, 17 of the 18 cotangents; 18, 19 are padding)
#pragma unroll
        for (int k = 0; k < ARP_ADJ_GROUP; k++) {
            const uint32_t q = min(qb + (uint32_t)k, q1 - 1u);
            sq[k] = a.sSid[q];  // (static: the entry's solid, no particle -> solid lookup)
            gq[k] = G[q];
            g2q[k] = G2[q];
        }
#pragma unroll
        for (int k = 0; k < ARP_ADJ_GROUP; k++) {
            const float4* F4 = reinterpret_cast<const float4*>(aa.Fbar + kFbarStride * (size_t)sq[k]);
#pragma unroll
            for (int r = 0; r < 4; r++) Fq[k][r] = F4[r];
            Ft[k] = reinterpret_cast<const float2*>(F4)[8];
        }
        // (every entry of the group is evaluated and a tail entry's sums are discarded by
        // selection, so that the group's loads stay ahead of the arithmetic)
#pragma unroll
        for (int k = 0; k < ARP_ADJ_GROUP; k++) {
            const bool live = qb + (uint32_t)k < q1;
            const float4 g = gq[k];
            const float gg[3] = {g.x, g.y, g.z};
            // own F: dx = x_q - x_s enters with weight V2 L_s^-1 gradW(dX)
            for (int r = 0; r < 3; r++) {
                const float t = g.w * (f1[3 * r] * gg[0] + f1[3 * r + 1] * gg[1] + f1[3 * r + 2] * gg[2]);
                const float td = g.w * (f1[9 + 3 * r] * gg[0] + f1[9 + 3 * r + 1] * gg[1] + f1[9 + 3 * r + 2] * gg[2]);
                xb[r] = live ? xb[r] - t : xb[r];
                vb[r] = live ? vb[r] - td : vb[r];
            }
            // the neighbour's F holds this particle with weight Vp L_s2^-1 gradW(X_s - X_s2)
            // (static: formed at upload)
            const float g2[3] = {g2q[k].x, g2q[k].y, g2q[k].z};
            const float fb2[20] = {Fq[k][0].x, Fq[k][0].y, Fq[k][0].z, Fq[k][0].w, Fq[k][1].x, Fq[k][1].y, Fq[k][1].z,
                                   Fq[k][1].w, Fq[k][2].x, Fq[k][2].y, Fq[k][2].z, Fq[k][2].w, Fq[k][3].x, Fq[k][3].y,
                                   Fq[k][3].z, Fq[k][3].w, Ft[k].x, Ft[k].y, 0.0f, 0.0f};
            for (int r = 0; r < 3; r++) {
                const float t = Vp * (fb2[3 * r] * g2[0] + fb2[3 * r + 1] * g2[1] + fb2[3 * r + 2] * g2[2]);
                const float td = Vp * (fb2[9 + 3 * r] * g2[0] + fb2[9 + 3 * r + 1] * g2[1] + fb2[9 + 3 * r + 2] * g2[2]);
                xb[r] = live ? xb[r] + t : xb[r];
                vb[r] = live ? vb[r] + td : vb[r];
            }
        }
    }
    float4* GP = reinterpret_cast<float4*>(aa.Gp) + i;
    float4* GV = reinterpret_cast<float4*>(aa.Gv) + i;
    float4* LP = reinterpret_cast<float4*>(aa.LamP) + i;
    float4* LV = reinterpret_cast<float4*>(aa.LamV) + i;
    float4 gp = *GP, gv = *GV, lp = *LP, lv = *LV;
    gp.x += xb[0]; gp.y += xb[1]; gp.z += xb[2];
    gv.x += vb[0]; gv.y += vb[1]; gv.z += vb[2];
    lp.x += xb[0]; lp.y += xb[1]; lp.z += xb[2];
    lv.x += vb[0]; lv.y += vb[1]; lv.z += vb[2];
    *GP = gp; *GV = gv; *LP = lp; *LV = lv;
}
