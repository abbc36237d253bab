// kernels.hip -- gfx950 kernels of the particle step (SPH fluid, neighbour grid, RK
// epilogue, std430 I/O).  Built with --genco into a code object that the host runtime
// loads with hipModuleLoadData and launches with hipModuleLaunchKernel.
//
// Reference semantics are cited per kernel (paths relative to the reference repository).
#include <hip/hip_runtime.h>

#include "common.h"

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
    unsigned long long active = __ballot(valid);
    const uint32_t lane = __lane_id();
    const unsigned long long lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    uint32_t rank = 0;
    while (active) {
        int leader = __ffsll(active) - 1;
        uint32_t lb = __shfl(bin, leader);
        unsigned long long m = __ballot(bin == lb) & active;
        uint32_t base = 0;
        if (lane == (uint32_t)leader) base = atomicAdd(&counters[lb], (uint32_t)__popcll(m));
        base = __shfl(base, leader);
        if ((m >> lane) & 1ull) rank = base + (uint32_t)__popcll(m & lt);
        active &= ~m;
    }
    return rank;
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

// EOS (fluid.rs:45-52)
__device__ __forceinline__ float eos_pressure(uint32_t eq, float density, float c, float d0) {
    switch (eq) {
        case 1: return c * c;
        case 2: return c * c * (density - d0);
        case 3: return (c * c * d0 / 7.0f) * (powf(density / d0, 7.0f) - 1.0f);
        default: return 0.0f;
    }
}

// p / (den*den) of a state, the per-particle half of the pressure term (fluid.rs:183-186)
__device__ __forceinline__ float pressure_q(const DMat& m, float den) {
    return eos_pressure(m.eq, den, m.c, m.target_den) / (den * den);
}

// contact_potential (fluid.rs:54-67)
__device__ __forceinline__ float contact_potential(uint32_t eq, float r, float strength, float r0) {
    switch (eq) {
        case 1: return r <= r0 ? strength : 0.0f;
        case 2: return strength * (r0 - r);
        case 3: return (strength * r / 7.0f) * (powf(r0 / r, 7.0f) - 1.0f);
        case 4: {
            float f = r0 / r;
            f *= f * f;
            f *= f;
            return strength * (f * f - 2.0f * f);
        }
        case 5: return (strength * r0 / 7.0f) * (powf(r / r0, 7.0f) - 1.0f);
        default: return 0.0f;
    }
}

// ------------------------------------------------------------------------------------
// neighbour grid: cell-row counting sort (replaces fill_buckets / reset_buckets,
// neighbor_list.rs:37-143)
// ------------------------------------------------------------------------------------

// Cell key of every particle + slot in its (y,z) row, reserved by wave-aggregated atomics.
extern "C" __global__ void __launch_bounds__(256) k_key(KeyArgs a) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    const bool in = i < a.n;
    uint32_t bin = 0, key = kInvalid;
    if (in) {
        float4 p = reinterpret_cast<const float4*>(a.posd)[i];
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

// Exclusive scan of the row counts (one 1024-thread block); lists the occupied rows and
// clears the counters for the next build.
extern "C" __global__ void __launch_bounds__(1024) k_scan(ScanArgs a) {
    __shared__ uint32_t sc[1024];
    __shared__ uint32_t sf[1024];
    const uint32_t T = 1024, t = threadIdx.x, nb = a.nbins;
    const uint32_t per = (nb + T - 1) / T;
    const uint32_t b0 = min(t * per, nb), b1 = min(b0 + per, nb);
    uint32_t s = 0, f = 0;
    for (uint32_t b = b0; b < b1; b++) {
        uint32_t c = a.binCount[b];
        s += c;
        f += (c > 0 && b + 1 < nb) ? 1u : 0u;
    }
    sc[t] = s;
    sf[t] = f;
    __syncthreads();
    for (uint32_t o = 1; o < T; o <<= 1) {
        uint32_t vs = t >= o ? sc[t - o] : 0, vf = t >= o ? sf[t - o] : 0;
        __syncthreads();
        sc[t] += vs;
        sf[t] += vf;
        __syncthreads();
    }
    uint32_t es = sc[t] - s, ef = sf[t] - f;
    for (uint32_t b = b0; b < b1; b++) {
        uint32_t c = a.binCount[b];
        a.binStart[b] = es;
        if (c > 0 && b + 1 < nb) a.activeRows[ef++] = b;
        es += c;
        a.binCount[b] = 0;
    }
    if (t == T - 1) {
        a.binStart[nb] = sc[T - 1];
        *a.numActive = sf[T - 1];
    }
}

__device__ __forceinline__ void move_particle(const ScatterArgs& a, uint32_t src, uint32_t dst) {
    reinterpret_cast<float4*>(a.dstPosd)[dst] = reinterpret_cast<const float4*>(a.srcPosd)[src];
    reinterpret_cast<float4*>(a.dstVelq)[dst] = reinterpret_cast<const float4*>(a.srcVelq)[src];
    a.dstMk[dst] = a.srcMk[src];
    uint32_t p = a.srcPid[src];
    a.dstPid[dst] = p;
    if (a.inv) a.inv[p] = dst;
}

// Scatter (key, source) into the reserved row slots; particles with non-finite positions
// go straight to their final place after all rows (neighbor_list.rs:86-93 drops them).
extern "C" __global__ void __launch_bounds__(256) k_scatter(ScatterArgs a) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= a.n) return;
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
        } else {
            a.perm[dst] = i | kFlag;
        }
    }
}

// One workgroup per occupied (y,z) row: LDS counting sort by x, then ascending particle id
// inside each cell (the order the 128-slot cap keeps), per-row cell offsets, cap flags,
// and the permutation / reorder of the state into cell-major storage.
extern "C" __global__ void __launch_bounds__(256) k_rowsort(RowSortArgs a) {
    __shared__ uint32_t cnt[kMaxSX + 1];
    __shared__ uint32_t off[kMaxSX + 1];
    __shared__ uint32_t ePid[kRowCap];
    __shared__ uint32_t eGrp[kRowCap];
    __shared__ uint16_t eX[kRowCap];
    __shared__ uint32_t wsum[4];
    const uint32_t tid = threadIdx.x;
    const uint32_t sx = a.grid.sub[0];
    const uint32_t nact = *a.numActive;
    for (uint32_t ri = blockIdx.x; ri < nact; ri += gridDim.x) {
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
            ro[x] = off[x];
            if (x < sx) cnt[x] = 0;  // reused as fill counters
        }
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
            } else {
                a.perm[dst] = src | flag;
            }
            if (flag && a.stats) atomicAdd(a.stats, 1u);
        }
        __syncthreads();
    }
}

// ------------------------------------------------------------------------------------
// fluid_forces (fluid.rs:95-251) with the RK stage epilogue fused in
// ------------------------------------------------------------------------------------
constexpr int kList = 16;  // within-h neighbours buffered per thread before a flush

struct PairAcc {
    float den, fx, fy, fz;
};

extern "C" __global__ void __launch_bounds__(256) k_fluid(FluidArgs a) {
    __shared__ uint32_t list[kList * 256];
    const uint32_t tid = threadIdx.x;
    const uint32_t t = blockIdx.x * 256u + tid;
    const bool act = t < a.n;
    const float4* X = reinterpret_cast<const float4*>(a.Xposd);
    const float4* XV = reinterpret_cast<const float4*>(a.Xvelq);
    const float4* B = reinterpret_cast<const float4*>(a.bposd);
    const float4* BV = reinterpret_cast<const float4*>(a.bvelq);
    const float h = a.phys.h, hh = h * h, kn = a.phys.k;

    uint32_t i = 0, mk1 = 0, nextBin = 0;
    bool haveKey = false, doKey = false;
    uint32_t nextKey = kInvalid;
    if (act) {
        const uint32_t pe = a.perm[t];
        i = pe & ~kFlag;
        const bool noNb = (pe & kFlag) != 0;
        const float4 xi = X[i];
        const float4 vi = XV[i];
        mk1 = a.mk[i];
        const uint32_t mat1 = mk1 & kMatMask;
        const bool elastic1 = (mk1 & kMkElastic) != 0;
        const DMat& M1 = a.mats[mat1];
        const float m1 = M1.mass, d1 = xi.w, q1 = vi.w;
        const float f1 = M1.visc, bf1 = M1.bulk_visc;
        PairAcc acc = {0.f, 0.f, 0.f, 0.f};

        auto pair = [&](uint32_t e) {
            const bool bnd = (e & kBnd) != 0;
            const uint32_t j = e & ~kBnd;
            float4 xj, vj;
            uint32_t mk2;
            if (bnd) {
                xj = B[j];
                vj = BV[j];
                mk2 = a.bmk[j];
            } else {
                xj = X[j];
                vj = XV[j];
                mk2 = a.mk[j];
            }
            const uint32_t mat2 = mk2 & kMatMask;
            const bool elastic2 = (mk2 & kMkElastic) != 0;
            const DMat& M2 = a.mats[mat2];
            const float m2 = M2.mass, d2 = xj.w, q2 = vj.w;
            const float rx = xj.x - xi.x, ry = xj.y - xi.y, rz = xj.z - xi.z;
            const float vx = vj.x - vi.x, vy = vj.y - vi.y, vz = vj.z - vi.z;
            const float rr = rx * rx + ry * ry + rz * rz;
            const float diff = fmaxf(hh - rr, 0.0f);
            const float gs = -6.0f * kn * diff * diff;
            const float gx = gs * rx, gy = gs * ry, gz = gs * rz;
            if (!bnd) acc.den += m2 * (vx * gx + vy * gy + vz * gz);
            const DInter it = a.inter[mat1 * a.phys.nm + mat2];
            const float cp = contact_potential(it.potential, sqrtf(rr), it.radius, it.strength);
            const float pf = m1 * m2 * (q1 + q2);
            acc.fx += pf * gx;
            acc.fy += pf * gy;
            acc.fz += pf * gz;
            if (cp != 0.0f) {
                const float cs = m2 * cp / d2;
                const float cx = cs * gx, cy = cs * gy, cz = cs * gz;
                if ((elastic1 || elastic2 || bnd) && mat1 != mat2) {
                    // friction along the tangential relative velocity (fluid.rs:191-199)
                    const float ix = rx / rr, iy = ry / rr, iz = rz / rr;
                    const float cfr = cx * rx + cy * ry + cz * rz;
                    const float vr = vx * rx + vy * ry + vz * rz;
                    const float nx = ix * cfr, ny = iy * cfr, nz = iz * cfr;
                    float tx = vx - ix * vr, ty = vy - iy * vr, tz = vz - iz * vr;
                    const float tl = sqrtf(tx * tx + ty * ty + tz * tz);
                    tx = tx / tl;
                    ty = ty / tl;
                    tz = tz / tl;
                    if (isfinite(tx) && isfinite(ty) && isfinite(tz)) {
                        const float fl = it.friction * sqrtf(nx * nx + ny * ny + nz * nz);
                        const float tl2 = sqrtf(tx * tx + ty * ty + tz * tz);
                        acc.fx += fl * (tx / tl2);
                        acc.fy += fl * (ty / tl2);
                        acc.fz += fl * (tz / tl2);
                    }
                }
                acc.fx += cx;
                acc.fy += cy;
                acc.fz += cz;
            }
            if (!bnd) {
                const float den_ = rr + a.phys.eps_hh;
                if (!elastic1) {
                    const float s = m1 * m2 * (f1 + M2.visc) * (rx * gx + ry * gy + rz * gz);
                    const float dd = d1 * d2 * den_;
                    acc.fx -= s * vx / dd;
                    acc.fy -= s * vy / dd;
                    acc.fz -= s * vz / dd;
                }
                const float ff = mat1 == mat2 ? bf1 + M2.bulk_visc : it.dampening;
                const float s = m1 * m2 * (ff * (vx * rx + vy * ry + vz * rz) / ((d1 + d2) * den_));
                acc.fx -= s * gx;
                acc.fy -= s * gy;
                acc.fz -= s * gz;
            }
        };

        if (!noNb) {
            uint32_t cx, cy, cz;
            cell_of(a.grid, xi.x, xi.y, xi.z, cx, cy, cz);
            const uint32_t sx = a.grid.sub[0], sy = a.grid.sub[1], sz = a.grid.sub[2];
            const uint32_t xl = cx > 0 ? cx - 1 : 0, xh = min(cx + 1, sx - 1);
            uint32_t nl = 0;
            auto flush = [&]() {
                for (uint32_t k = 0; k < nl; k++) pair(list[k * 256 + tid]);
                nl = 0;
            };
            auto push = [&](uint32_t e) {
                list[nl * 256 + tid] = e;
                if (++nl == (uint32_t)kList) flush();
            };
            for (int dz = -1; dz <= 1; dz++) {
                const int nz = (int)cz + dz;
                if (nz < 0 || nz >= (int)sz) continue;
                for (int dy = -1; dy <= 1; dy++) {
                    const int ny = (int)cy + dy;
                    if (ny < 0 || ny >= (int)sy) continue;
                    const uint32_t row = (uint32_t)ny + sy * (uint32_t)nz;
                    const uint32_t s0 = a.binStart[row], s1 = a.binStart[row + 1];
                    if (s1 > s0) {
                        const uint32_t* ro = a.rowOff + (size_t)row * a.grid.rowstride;
                        const uint32_t qa = s0 + ro[xl], qb = s0 + ro[xh + 1];
                        for (uint32_t q = qa; q < qb; q++) {
                            const uint32_t pe2 = a.perm[q];
                            if (pe2 & kFlag) continue;
                            const float4 xj = X[pe2];
                            const float rx = xj.x - xi.x, ry = xj.y - xi.y, rz = xj.z - xi.z;
                            if (hh - (rx * rx + ry * ry + rz * rz) > 0.0f) push(pe2);
                        }
                    }
                    const size_t cb = (size_t)row * sx;
                    const uint32_t ba = a.bP[cb + xl], bb = a.bP[cb + xh + 1];
                    for (uint32_t q = ba; q < bb; q++) {
                        const float4 xj = B[q];
                        const float rx = xj.x - xi.x, ry = xj.y - xi.y, rz = xj.z - xi.z;
                        if (hh - (rx * rx + ry * ry + rz * rz) > 0.0f) push(q | kBnd);
                    }
                }
            }
            flush();
        }

        // derivative (fluid.rs:243-247): den' = sum, vel' = F/den + (0,-g,0), pos' = vel
        const float kPx = vi.x, kPy = vi.y, kPz = vi.z, kPw = acc.den;
        const float kVx = acc.fx / d1, kVy = acc.fy / d1 - a.phys.g, kVz = acc.fz / d1;

        if (a.mode == 2 || (a.skip_elastic && elastic1)) {
            reinterpret_cast<float4*>(a.KpOut)[i] = make_float4(kPx, kPy, kPz, kPw);
            reinterpret_cast<float4*>(a.KvOut)[i] = make_float4(kVx, kVy, kVz, 0.0f);
        } else {
            const float4 yp = reinterpret_cast<const float4*>(a.Yposd)[i];
            const float4 yv = reinterpret_cast<const float4*>(a.Yvelq)[i];
            float4 ap = make_float4(0.f, 0.f, 0.f, 0.f), av = ap;
            if (a.AccP) {
                ap = reinterpret_cast<const float4*>(a.AccP)[i];
                av = reinterpret_cast<const float4*>(a.AccV)[i];
            }
            const float b = a.b_cur;
            ap.x += b * kPx; ap.y += b * kPy; ap.z += b * kPz; ap.w += b * kPw;
            av.x += b * kVx; av.y += b * kVy; av.z += b * kVz;
            float4 op, ov;
            if (a.last) {
                op = make_float4(yp.x + ap.x, yp.y + ap.y, yp.z + ap.z, yp.w + ap.w);
                ov = make_float4(yv.x + av.x, yv.y + av.y, yv.z + av.z, 0.f);
            } else {
                const float c = a.a_next;
                op = make_float4(yp.x + c * kPx, yp.y + c * kPy, yp.z + c * kPz, yp.w + c * kPw);
                ov = make_float4(yv.x + c * kVx, yv.y + c * kVy, yv.z + c * kVz, 0.f);
                for (int kk = 0; kk < a.nk; kk++) {
                    const float4 p2 = reinterpret_cast<const float4*>(a.Kp[kk])[i];
                    const float4 v2 = reinterpret_cast<const float4*>(a.Kv[kk])[i];
                    const float ck = a.kc[kk];
                    op.x += ck * p2.x; op.y += ck * p2.y; op.z += ck * p2.z; op.w += ck * p2.w;
                    ov.x += ck * v2.x; ov.y += ck * v2.y; ov.z += ck * v2.z;
                }
                if (a.AccPOut) {
                    reinterpret_cast<float4*>(a.AccPOut)[i] = ap;
                    reinterpret_cast<float4*>(a.AccVOut)[i] = av;
                }
            }
            if (a.KpOut) {
                reinterpret_cast<float4*>(a.KpOut)[i] = make_float4(kPx, kPy, kPz, kPw);
                reinterpret_cast<float4*>(a.KvOut)[i] = make_float4(kVx, kVy, kVz, 0.0f);
            }
            ov.w = pressure_q(M1, op.w);
            reinterpret_cast<float4*>(a.Oposd)[i] = op;
            reinterpret_cast<float4*>(a.Ovelq)[i] = ov;
            doKey = a.keys != nullptr;
            if (doKey) {
                uint32_t cx, cy, cz;
                haveKey = true;
                if (cell_of(a.grid, op.x, op.y, op.z, cx, cy, cz)) {
                    const uint32_t row = cy + a.grid.sub[1] * cz;
                    nextKey = cx + a.grid.sub[0] * row;
                    nextBin = row;
                } else {
                    nextBin = a.grid.nrows;
                }
            }
        }
    }
    if (a.keys) {
        const uint32_t rank = agg_rank(a.binCount, nextBin, haveKey);
        if (haveKey) {
            a.keys[i] = nextKey;
            a.ranks[i] = rank;
        }
    }
}

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
