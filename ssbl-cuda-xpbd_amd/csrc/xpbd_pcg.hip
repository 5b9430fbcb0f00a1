// xpbd_pcg.hip -- the hard-stretch path on gfx950: block-Jacobi PCG on the
// edge graph, then the polish / tiny-edge / cap passes.
//
// Reference (native/src/ssbl_xpbd_cuda_abi41.cu, A41:<line>): PCG kernels
// A41:3032-3464 and driver A41:7626-7759 (8 iterations, solution written back
// to the positions every iteration), polish A41:1982-2112, tiny-edge pass
// A41:2399-2494, caps A41:2249-2362 / 2496-2601 / 2603-2708, sequences
// A41:7761-8027, schedule A41:9183-9201 and 9344-9350.
//
// Every edge-accumulating kernel of the reference (system build, polish and
// tiny-edge averages, uncoloured caps) sums with float atomics; here each
// vertex gathers its incident edges in ascending edge order.  The coloured cap
// passes keep the reference's one-launch-per-colour form (race free for a
// valid colouring).  The PCG inner products are summed exactly: each float
// term becomes a 2^-96 fixed-point __int128, partial sums per workgroup, one
// final 128-bit sum rounded to double -- the result does not depend on the
// order of the additions.
//
// Batching: every kernel reads the cloth it works on from the step's view
// table (the tile's or segment's cloth index), so one launch covers all the
// cloths of a group; a cloth's workgroup partials sit at its own tile
// indices, and the reduction runs one workgroup per cloth.
#include "xpbd_pcg.hpp"

#include "xpbd_math.hpp"
#include "xpbd_timing.hpp"

#include <algorithm>
#include <mutex>
#include <vector>

namespace ssbl {

// The static scratch of the allocation-free step (ssbl_ext_step_solver_ex_noalloc):
// one solve's vectors, blocks, partials and scalars.  In host code `scratch`
// names a host-side stand-in, not the device array: kernels only ever get the
// device address looked up below.
static __device__ __align__(256) unsigned char scratch[kPcgScratchBytes];

namespace {
constexpr int kScratchDevices = 64;
std::mutex g_scratch_mutex;
void* g_scratch_base[kScratchDevices] = {};
}  // namespace

bool pcg_static_scratch(void** base, size_t* bytes) {
    int dev = -1;
    if (!hip_ok(hipGetDevice(&dev), "current device")) return false;
    if (dev < 0 || dev >= kScratchDevices) return fail("device index beyond the static scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_base[dev]) {
        void* p = nullptr;
        if (!hip_ok(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "static scratch address") || !p) return false;
        g_scratch_base[dev] = p;
    }
    *base = g_scratch_base[dev];
    *bytes = kPcgScratchBytes;
    return true;
}

namespace {

constexpr double kFix = 79228162514264337593543950336.0;   // 2^96
constexpr int kPcgIterations = 8;                           // A41:152

struct Sym { float m11, m12, m13, m22, m23, m33; };

// Incidence slots whose loads a gather issues together.
constexpr int kPcgChunk = 4;
static_assert(kEllMax % kPcgChunk == 0, "the chunk must divide the slot count");

__device__ __forceinline__ Sym load_sym(const float* p) { return Sym{p[0], p[1], p[2], p[3], p[4], p[5]}; }
__device__ __forceinline__ void store_sym(float* p, Sym s) { p[0] = s.m11; p[1] = s.m12; p[2] = s.m13; p[3] = s.m22; p[4] = s.m23; p[5] = s.m33; }

// A41:1193-1199
__device__ __forceinline__ F3 sym_madd(Sym m, F3 v, F3 acc) {
    return f3(__builtin_fmaf(m.m13, v.z, __builtin_fmaf(m.m12, v.y, __builtin_fmaf(m.m11, v.x, acc.x))),
              __builtin_fmaf(m.m23, v.z, __builtin_fmaf(m.m22, v.y, __builtin_fmaf(m.m12, v.x, acc.y))),
              __builtin_fmaf(m.m33, v.z, __builtin_fmaf(m.m23, v.y, __builtin_fmaf(m.m13, v.x, acc.z))));
}

// A41:1158-1191
__device__ inline Sym sym_inverse(Sym m, unsigned long long* guard) {
    const float a = m.m11, b = m.m12, c = m.m13, d = m.m22, e = m.m23, f = m.m33;
    const float df_ee = __builtin_fmaf(d, f, -e * e), bf_ce = __builtin_fmaf(b, f, -c * e), be_cd = __builtin_fmaf(b, e, -c * d);
    const float det = __builtin_fmaf(c, be_cd, __builtin_fmaf(-b, bf_ce, a * df_ee));
    if (isfinite(det) && fabsf(det) > 1.0e-12f) {
        const float inv = 1.0f / det;
        return Sym{df_ee * inv, __builtin_fmaf(c, e, -b * f) * inv, be_cd * inv, __builtin_fmaf(a, f, -c * c) * inv,
                   __builtin_fmaf(b, c, -a * e) * inv, __builtin_fmaf(a, d, -b * b) * inv};
    }
    atomicAdd(guard, 1ull);
    return Sym{1.0f / fmaxf(fabsf(a), 1.0e-4f), 0.0f, 0.0f, 1.0f / fmaxf(fabsf(d), 1.0e-4f), 0.0f, 1.0f / fmaxf(fabsf(f), 1.0e-4f)};
}

__device__ __forceinline__ __int128 to_fixed(float t) { return static_cast<__int128>(rint(static_cast<double>(t) * kFix)); }

// Workgroup maximum of non-negative floats (compared as bit patterns, which
// order them), returned to every thread.  Every thread must call it.
__device__ inline float block_max_nonneg(float v) {
    __shared__ int wave_m[kTile / 64];
    int m = __float_as_int(fmaxf(v, 0.0f));
    for (int d = 32; d > 0; d >>= 1) m = max(m, __shfl_xor(m, d, 64));
    if ((threadIdx.x & 63u) == 0u) wave_m[threadIdx.x >> 6] = m;
    __syncthreads();
    int r = wave_m[0];
    for (int w = 1; w < kTile / 64; ++w) r = max(r, wave_m[w]);
    return __int_as_float(r);
}

// Workgroup sum of one fixed-point term per thread -> partials[slot].
__device__ inline void block_partial(__int128 v, unsigned long long* partials, int slot) {
    __shared__ unsigned long long lo[kTile], hi[kTile];
    lo[threadIdx.x] = static_cast<unsigned long long>(v);
    hi[threadIdx.x] = static_cast<unsigned long long>(v >> 64);
    __syncthreads();
    for (int s = kTile / 2; s > 0; s >>= 1) {
        if (static_cast<int>(threadIdx.x) < s) {
            const __int128 a = (static_cast<__int128>(hi[threadIdx.x]) << 64) | lo[threadIdx.x];
            const __int128 b = (static_cast<__int128>(hi[threadIdx.x + s]) << 64) | lo[threadIdx.x + s];
            const __int128 c = a + b;
            lo[threadIdx.x] = static_cast<unsigned long long>(c);
            hi[threadIdx.x] = static_cast<unsigned long long>(c >> 64);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        partials[2 * slot] = lo[0];
        partials[2 * slot + 1] = hi[0];
    }
}

__device__ __forceinline__ int tile_of_block(int ntiles) {
    const int b = static_cast<int>(blockIdx.x);
    const int xcd = b & 7, local = b >> 3;
    const int per = ntiles >> 3, extra = ntiles & 7;
    return xcd * per + min(xcd, extra) + local;
}

__device__ __forceinline__ float4* cur(const ClothView& c, int parity) { return parity ? c.pos1 : c.pos0; }

// A41:3032-3069 reset + A41:3085-3186 system build (gathered per vertex in
// ascending edge order) + A41:3188-3226 preconditioner.
__global__ __launch_bounds__(kTile) void k_pcg_build(const ClothView* __restrict__ views, const Tile* __restrict__ tiles, int ntiles,
                                                     int parity) {
    const Tile tile = tiles[tile_of_block(ntiles)];
    const ClothView& c = views[tile.cloth];
    const PcgView& pv = c.pcg;
    const int v = tile.begin + static_cast<int>(threadIdx.x);
    __int128 term = 0;
    if (v < tile.end) {
        const float4* P = cur(c, parity);
        const float4 p4 = P[v];
        const float w = p4.w;
        const bool pinned = w == 0.0f;
        Sym d{1.0f, 0.0f, 0.0f, 1.0f, 0.0f, 1.0f};
        F3 rhs = f3(0.0f, 0.0f, 0.0f);
        auto incidence = [&](int2 a, float4 o4, size_t slot) {
            const float rest = __int_as_float(a.y);
            const float wo = o4.w;
            Sym off{0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
            const float wsum = w + wo;
            bool use = isfinite(wsum) && wsum > 0.0f && isfinite(rest) && rest > kEps;
            const F3 pv3 = f3(p4), po3 = f3(o4);
            use = use && finite3(pv3) && finite3(po3);
            F3 delta = po3 - pv3;
            float len = 0.0f, err = 0.0f;
            if (use) {
                const float lsq = dot(delta, delta);
                use = isfinite(lsq) && lsq > kEps;
                if (use) {
                    len = sqrtf(lsq);
                    err = len - rest;
                    use = isfinite(err) && fabsf(err) > 1.0e-6f && pv.strength > 0.0f;
                }
            }
            if (use) {
                float limit = fmaxf(1.0e-5f, fminf(fmaxf(rest, pv.thick) * 0.5f, 0.25f));   // A41:1070-1089
                if (pv.hard) limit = fminf(limit, fmaxf(1.0e-5f, fminf(fmaxf(rest * 0.18f, pv.thick * 0.06f), 0.0035f)));
                float proj = err * pv.strength;
                proj = fmaxf(-limit, fminf(proj, limit));
                if (isfinite(proj) && fabsf(proj) > 1.0e-7f) {
                    const F3 nrm = mul(delta, 1.0f / len);
                    const F3 desired = mul(nrm, proj / wsum);
                    if (w > 0.0f) rhs = rhs + mul(desired, w);
                    const float k_s = fmaxf(pv.strength, 1.0e-3f);
                    const Sym blk{nrm.x * nrm.x * k_s, nrm.x * nrm.y * k_s, nrm.x * nrm.z * k_s,
                                  nrm.y * nrm.y * k_s, nrm.y * nrm.z * k_s, nrm.z * nrm.z * k_s};
                    if (w > 0.0f) {
                        d.m11 += blk.m11; d.m12 += blk.m12; d.m13 += blk.m13; d.m22 += blk.m22; d.m23 += blk.m23; d.m33 += blk.m33;
                    }
                    if (w > 0.0f && wo > 0.0f) off = Sym{-blk.m11, -blk.m12, -blk.m13, -blk.m22, -blk.m23, -blk.m33};
                }
            }
            store_sym(pv.off + 6 * slot, off);
        };
        if (pv.ell_off) {
            // slot-major incidences, each chunk's loads issued together
#pragma unroll
            for (int j0 = 0; j0 < kEllMax; j0 += kPcgChunk) {
                if (j0 >= c.ell_width) break;
                int2 a[kPcgChunk];
                float4 o4[kPcgChunk];
#pragma unroll
                for (int j = 0; j < kPcgChunk; ++j) {
                    a[j] = j0 + j < c.ell_width ? c.ell[static_cast<size_t>(j0 + j) * c.ell_pitch + v] : make_int2(-1, 0);
                    o4[j] = P[a[j].x != -1 ? (a[j].x & 0x7fffffff) : v];
                }
#pragma unroll
                for (int j = 0; j < kPcgChunk; ++j)
                    if (a[j].x != -1) incidence(a[j], o4[j], static_cast<size_t>(j0 + j) * c.ell_pitch + v);
            }
        } else {
            const int b = c.adj_off[v], e = c.adj_off[v + 1];
            for (int k = b; k < e; ++k) {
                const int2 a = c.adj[k];
                incidence(a, P[a.x & 0x7fffffff], static_cast<size_t>(k));
            }
        }
        store_sym(pv.diag + 6 * static_cast<size_t>(v), d);
        const F3 zero = f3(0.0f, 0.0f, 0.0f);
        pv.sol[v] = with_w(zero, 0.0f);
        pv.adir[v] = with_w(zero, 0.0f);
        if (pinned) {
            pv.rhs[v] = pv.res[v] = pv.z[v] = pv.dir[v] = with_w(zero, 0.0f);
            store_sym(pv.pinv + 6 * static_cast<size_t>(v), Sym{1.0f, 0.0f, 0.0f, 1.0f, 0.0f, 1.0f});
        } else {
            const Sym inv = sym_inverse(d, pv.guard);
            const F3 z = sym_madd(inv, rhs, zero);
            store_sym(pv.pinv + 6 * static_cast<size_t>(v), inv);
            pv.rhs[v] = pv.res[v] = with_w(rhs, 0.0f);
            pv.z[v] = pv.dir[v] = with_w(z, 0.0f);
            term = to_fixed(dot(rhs, z));
        }
    }
    block_partial(term, pv.partials, tile.begin / kTile);
}

// Final sum of one cloth's workgroup partials, a workgroup per cloth of the
// group (members[blockIdx.x]).  mode 0: initial rz; 1: dAd; 2: rz_next, then
// the search-direction factor and the advance (A41:3421-3464) when another
// iteration follows.
constexpr int kReduceBlock = 256;
__global__ __launch_bounds__(kReduceBlock) void k_pcg_reduce(const ClothView* __restrict__ views, const int* __restrict__ members,
                                                             int mode, int more) {
    const PcgView& pv = views[members[blockIdx.x]].pcg;
    const int nparts = pv.ntiles;
    __shared__ unsigned long long lo[kReduceBlock], hi[kReduceBlock];
    __shared__ int top[kReduceBlock];
    __int128 acc = 0;
    int m = 0;   // mode 2 (after an update): the largest step over the tiles, as bits
    for (int i = static_cast<int>(threadIdx.x); i < nparts; i += kReduceBlock) {
        acc += (static_cast<__int128>(pv.partials[2 * i + 1]) << 64) | pv.partials[2 * i];
        if (mode == 2) m = max(m, __float_as_int(pv.tile_max[i]));
    }
    top[threadIdx.x] = m;
    lo[threadIdx.x] = static_cast<unsigned long long>(acc);
    hi[threadIdx.x] = static_cast<unsigned long long>(acc >> 64);
    __syncthreads();
    for (int s = kReduceBlock / 2; s > 0; s >>= 1) {
        if (static_cast<int>(threadIdx.x) < s) {
            const __int128 a = (static_cast<__int128>(hi[threadIdx.x]) << 64) | lo[threadIdx.x];
            const __int128 b = (static_cast<__int128>(hi[threadIdx.x + s]) << 64) | lo[threadIdx.x + s];
            const __int128 c = a + b;
            lo[threadIdx.x] = static_cast<unsigned long long>(c);
            hi[threadIdx.x] = static_cast<unsigned long long>(c >> 64);
            top[threadIdx.x] = max(top[threadIdx.x], top[threadIdx.x + s]);
        }
        __syncthreads();
    }
    if (threadIdx.x != 0) return;
    if (mode == 2 && top[0] > __float_as_int(*pv.max_delta)) *pv.max_delta = __int_as_float(top[0]);
    const __int128 total = (static_cast<__int128>(hi[0]) << 64) | lo[0];
    const double value = static_cast<double>(total) * (1.0 / kFix);
    PcgScalars& s = *pv.sc;
    if (mode == 0) {
        s.rz = value; s.initial = value; s.dad = 0.0; s.rznext = 0.0; s.status = 0; s.dir_ok = 0;
        s.runs += 1;
        return;
    }
    if (mode == 1) { s.dad += value; return; }
    s.rznext += value;
    s.dir_ok = 0;
    if (!more || s.status != 0) return;
    if (isfinite(s.rz) && isfinite(s.rznext) && s.rz > 1.0e-14 && s.rznext >= 0.0) {
        const float beta = static_cast<float>(s.rznext / s.rz);
        if (isfinite(beta)) { s.beta = beta; s.dir_ok = 1; }
    }
    if (isfinite(s.rznext) && s.rznext >= 0.0) { s.rz = s.rznext; s.dad = 0.0; s.rznext = 0.0; }
}

// A41:3228-3267: A d (block diagonal + CSR off-diagonal blocks).
__global__ __launch_bounds__(kTile) void k_pcg_ad(const ClothView* __restrict__ views, const Tile* __restrict__ tiles, int ntiles,
                                                  int parity) {
    const Tile tile = tiles[tile_of_block(ntiles)];
    const ClothView& c = views[tile.cloth];
    const PcgView& pv = c.pcg;
    const int v = tile.begin + static_cast<int>(threadIdx.x);
    __int128 term = 0;
    const bool ok = pv.sc->status == 0;
    if (v < tile.end && ok) {
        if (cur(c, parity)[v].w == 0.0f) {
            pv.adir[v] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        } else {
            const F3 d = f3(pv.dir[v]);
            F3 q = sym_madd(load_sym(pv.diag + 6 * static_cast<size_t>(v)), d, f3(0.0f, 0.0f, 0.0f));
            if (pv.ell_off) {
                // slot-major incidences and blocks, each chunk's loads issued
                // together; the sum runs in incidence order as below
#pragma unroll
                for (int j0 = 0; j0 < kEllMax; j0 += kPcgChunk) {
                    if (j0 >= c.ell_width) break;
                    int x[kPcgChunk];
                    float4 dj[kPcgChunk];
                    Sym m[kPcgChunk];
#pragma unroll
                    for (int j = 0; j < kPcgChunk; ++j) {
                        const size_t slot = static_cast<size_t>(j0 + j) * c.ell_pitch + v;
                        x[j] = j0 + j < c.ell_width ? c.ell[slot].x : -1;
                        dj[j] = pv.dir[x[j] != -1 ? (x[j] & 0x7fffffff) : v];
                        m[j] = load_sym(pv.off + 6 * (x[j] != -1 ? slot : static_cast<size_t>(v)));
                    }
#pragma unroll
                    for (int j = 0; j < kPcgChunk; ++j)
                        if (x[j] != -1) q = sym_madd(m[j], f3(dj[j]), q);
                }
            } else {
                const int b = c.adj_off[v], e = c.adj_off[v + 1];
                for (int k = b; k < e; ++k) {
                    const int o = c.adj[k].x & 0x7fffffff;
                    q = sym_madd(load_sym(pv.off + 6 * static_cast<size_t>(k)), f3(pv.dir[o]), q);
                }
            }
            pv.adir[v] = with_w(q, 0.0f);
            term = to_fixed(dot(d, q));
        }
    }
    block_partial(term, pv.partials, tile.begin / kTile);
}

// A41:3320-3405 (device alpha) with the write-back of the step to the positions.
__global__ __launch_bounds__(kTile) void k_pcg_update(const ClothView* __restrict__ views, const Tile* __restrict__ tiles, int ntiles,
                                                      int parity) {
    const Tile tile = tiles[tile_of_block(ntiles)];
    const ClothView& c = views[tile.cloth];
    const PcgView& pv = c.pcg;
    const int v = tile.begin + static_cast<int>(threadIdx.x);
    const PcgScalars& s = *pv.sc;
    int status = 0;
    float alpha = 0.0f;
    if (!isfinite(s.rz) || s.rz < 0.0) status = 1;
    else if (s.rz <= 1.0e-14) status = 2;
    else if (!isfinite(s.dad) || s.dad <= 1.0e-14) status = 3;
    else {
        alpha = static_cast<float>(s.rz / s.dad);
        if (!isfinite(alpha)) status = 4;
    }
    __int128 term = 0;
    bool clamped = false;
    float moved = 0.0f;
    if (status == 0 && v < tile.end) {
        float4* P = cur(c, parity);
        const float4 p4 = P[v];
        const F3 zero = f3(0.0f, 0.0f, 0.0f);
        if (p4.w == 0.0f) {
            pv.sol[v] = pv.res[v] = pv.z[v] = with_w(zero, 0.0f);
        } else {
            const F3 d = f3(pv.dir[v]), ad = f3(pv.adir[v]), so = f3(pv.sol[v]), re = f3(pv.res[v]);
            F3 sol = f3(__builtin_fmaf(d.x, alpha, so.x), __builtin_fmaf(d.y, alpha, so.y), __builtin_fmaf(d.z, alpha, so.z));
            F3 res = f3(__builtin_fmaf(ad.x, -alpha, re.x), __builtin_fmaf(ad.y, -alpha, re.y), __builtin_fmaf(ad.z, -alpha, re.z));
            if (!finite3(sol)) sol = zero;
            if (!finite3(res)) res = zero;
            const F3 z = sym_madd(load_sym(pv.pinv + 6 * static_cast<size_t>(v)), res, zero);
            pv.sol[v] = with_w(sol, 0.0f);
            pv.res[v] = with_w(res, 0.0f);
            pv.z[v] = with_w(z, 0.0f);
            term = to_fixed(dot(res, z));
            F3 delta = mul(d, alpha);
            const float l = len3(delta);
            if (l > pv.wb_limit) {
                delta = mul(delta, pv.wb_limit / fmaxf(l, kEps));
                clamped = true;
            }
            P[v] = with_w(f3(p4) + delta, p4.w);
            const float4 q = c.prev[v];
            c.prev[v] = with_w(f3(q) + mul(delta, 0.08f), 0.0f);
            const float dl = len3(delta);
            if (dl > 0.0f) moved = dl;
        }
    }
    // the clamp count with one atomic per wave; the largest step per
    // workgroup, merged into the cloth's maximum by the reduction that
    // follows (one atomic per wave on the cloth's one address serialises
    // on large cloths)
    wave_count(c.counts, kCountHardFallbacks, clamped);
    const float block_moved = block_max_nonneg(moved);
    if (threadIdx.x == 0) pv.tile_max[tile.begin / kTile] = block_moved;
    // every workgroup of the cloth derives the same status; its first tile's records it
    if (tile.begin == 0 && threadIdx.x == 0) {
        pv.sc->status = status;
        if (status != 0 && status != 2) atomicAdd(pv.guard, 1ull);
    }
    block_partial(term, pv.partials, tile.begin / kTile);
}

// A41:3421-3450 with the factor computed by k_pcg_reduce.
__global__ __launch_bounds__(kTile) void k_pcg_dir(const ClothView* __restrict__ views, const Tile* __restrict__ tiles, int ntiles,
                                                   int parity) {
    const Tile tile = tiles[tile_of_block(ntiles)];
    const ClothView& c = views[tile.cloth];
    const PcgView& pv = c.pcg;
    if (!pv.sc->dir_ok) return;
    const int v = tile.begin + static_cast<int>(threadIdx.x);
    if (v >= tile.end) return;
    if (cur(c, parity)[v].w == 0.0f) { pv.dir[v] = make_float4(0.0f, 0.0f, 0.0f, 0.0f); return; }
    const float beta = pv.sc->beta;
    const F3 d = f3(pv.dir[v]), z = f3(pv.z[v]);
    pv.dir[v] = with_w(f3(__builtin_fmaf(d.x, beta, z.x), __builtin_fmaf(d.y, beta, z.y), __builtin_fmaf(d.z, beta, z.z)), 0.0f);
}

// ------------------------------------------------- polish / tiny / caps
struct EdgeOut { F3 d, q; bool apply; };

// One edge seen from vertex v ("this" end) with the other end o; the kinds
// and formulas are those of the reference kernels cited in the file header:
// 0 normal cap, 1 tiny-edge cap, 2 extreme cap, 3 polish contribution,
// 4 tiny-edge contribution.  The other end's delta is the same expression
// with the ends swapped, so a gather from either end reproduces the edge
// kernel's per-end result exactly.
__device__ inline bool edge_kind(int kind, F3 pi, F3 pj, F3 qi, F3 qj, float wi, float wj, float rest, float thick, float strength,
                                 float dt, EdgeOut* out) {
    out->apply = false;
    const bool tiny_kind = kind == 1 || kind == 4;
    if (!isfinite(rest) || rest <= kEps) return false;
    if (kind != 3 && (tiny_kind ? rest > 7.5e-4f : rest <= 7.5e-4f)) return false;
    const float w = wi + wj;
    if (!isfinite(w) || w <= 0.0f) return false;
    if (!finite3(pi) || !finite3(pj)) return false;
    if (kind != 4 && (!finite3(qi) || !finite3(qj))) return false;
    const F3 edge = pj - pi;
    const float lsq = dot(edge, edge);
    if (!isfinite(lsq) || lsq <= 1.0e-14f) return false;
    const float len = sqrtf(lsq);
    const bool high = strength >= 0.95f;
    if (kind >= 3) {
        const float target = kind == 3 ? rest * (high ? 1.08f : 1.18f) : rest * 1.06f;
        const float excess = len - target;
        if (!isfinite(excess) || excess <= 0.0f) return false;
        const F3 corr = mul(edge, (kind == 3 ? -0.90f : -1.0f) * excess / fmaxf(w * len, kEps));
        if (!finite3(corr) || len3(corr) <= 1.0e-20f) return false;
        const float md = kind == 3 ? fmaxf(5.0e-4f, fminf(fmaxf(rest * 2.0f, thick * 0.25f), 0.008f))
                                   : fminf(fmaxf(rest * 3.0f, 2.5e-4f), 8.0e-4f);
        if (wi > 0.0f) {
            const F3 d = limit_delta(mul(corr, -wi), md);
            if (finite3(d) && len3(d) > 0.0f) { out->d = d; out->apply = true; }
        }
        return true;
    }
    float start, target, md, sync;
    if (kind == 0) {
        start = rest * (high ? 1.12f : 1.32f); target = rest * (high ? 1.04f : 1.22f);
        md = fminf(fmaxf(rest * 0.60f, thick * 0.10f), 0.010f); sync = 0.08f;
    } else if (kind == 1) {
        start = rest * 1.35f; target = rest * 1.12f; md = fminf(fmaxf(rest * 6.0f, 5.0e-4f), 1.5e-3f); sync = 1.0f;
    } else {
        start = rest * 1.24f; target = rest * 1.10f; md = fminf(fmaxf(rest * 1.25f, thick * 0.10f), 0.006f); sync = 0.30f;
    }
    if (!isfinite(len) || len <= start) return false;
    const float excess = len - target;
    if (!isfinite(excess) || excess <= 0.0f) return false;
    const F3 corr = mul(edge, -excess / fmaxf(w * len, kEps));
    if (!finite3(corr) || len3(corr) <= 1.0e-20f) return false;
    F3 pd = f3(0.0f, 0.0f, 0.0f);
    const F3 nrm = mul(edge, 1.0f / fmaxf(len, kEps));
    const float sdt = kind == 0 ? fmaxf(dt, kEps) : dt;
    if (isfinite(sdt) && sdt > kEps) {
        const float inv = 1.0f / sdt;
        const F3 vi = mul(pi - qi, inv), vj = mul(pj - qj, inv);
        const float sep = dot(vj - vi, nrm);
        if (isfinite(sep) && sep > 0.0f) {
            const float trim = sep * sdt * 1.0f / w;
            if (wi > 0.0f) pd = pd + mul(nrm, -wi * trim);
        }
    }
    if (wi > 0.0f) {
        const F3 d = limit_delta(mul(corr, -wi), md);
        if (finite3(d) && len3(d) > 0.0f) { out->d = d; out->q = pd + mul(d, sync); out->apply = true; }
    }
    return true;
}

// Jacobi gather of one kind over every incident edge (src -> dst): caps add
// each edge's result in ascending edge order; the polish / tiny-edge kinds
// average their contributions (A41:2082-2112, 2468-2494).
__global__ __launch_bounds__(kTile) void k_edge_gather(const ClothView* __restrict__ views, const Tile* __restrict__ tiles, int ntiles,
                                                       int src_parity, int kind) {
    const Tile tile = tiles[tile_of_block(ntiles)];
    const int v = tile.begin + static_cast<int>(threadIdx.x);
    if (v >= tile.end) return;
    const ClothView& c = views[tile.cloth];
    const PcgView& pv = c.pcg;
    const float dt = c.dt_sub;
    const float4* S = src_parity ? c.pos1 : c.pos0;
    float4* D = src_parity ? c.pos0 : c.pos1;
    if (kind == -2) { pv.qsnap[v] = c.prev[v]; return; }   // snapshot of the previous positions
    const float4 p4 = S[v];
    if (kind < 0) { D[v] = p4; return; }   // plain copy (parity realignment)
    const F3 p0 = f3(p4), q0 = f3(c.prev[v]);
    // kinds 0-2 use the neighbours' previous positions: from the snapshot,
    // since this launch rewrites them (kinds 3-4 only test them for being
    // finite, which the rewrite keeps)
    const float4* Qn = kind <= 2 && pv.qsnap ? pv.qsnap : c.prev;
    const float w = p4.w;
    F3 p = p0, q = q0, sum = f3(0.0f, 0.0f, 0.0f);
    int cnt = 0;
    auto incidence = [&](int2 a, float4 o4, float4 qo) {
        EdgeOut out;
        if (!edge_kind(kind, p0, f3(o4), q0, f3(qo), w, o4.w, __int_as_float(a.y), pv.thick, pv.strength_cfg, dt, &out)) return;
        if (!out.apply) return;
        if (kind >= 3) { sum = sum + out.d; ++cnt; }
        else { p = p + out.d; q = q + out.q; }
    };
    if (pv.ell_off) {
        // slot-major incidences (the same order), each chunk's loads issued together
#pragma unroll
        for (int j0 = 0; j0 < kEllMax; j0 += kPcgChunk) {
            if (j0 >= c.ell_width) break;
            int2 a[kPcgChunk];
            float4 o4[kPcgChunk], qo[kPcgChunk];
#pragma unroll
            for (int j = 0; j < kPcgChunk; ++j) {
                a[j] = j0 + j < c.ell_width ? c.ell[static_cast<size_t>(j0 + j) * c.ell_pitch + v] : make_int2(-1, 0);
                const int o = a[j].x != -1 ? (a[j].x & 0x7fffffff) : v;
                o4[j] = S[o];
                qo[j] = Qn[o];
            }
#pragma unroll
            for (int j = 0; j < kPcgChunk; ++j)
                if (a[j].x != -1) incidence(a[j], o4[j], qo[j]);
        }
    } else {
        const int b = c.adj_off[v], e = c.adj_off[v + 1];
        for (int k = b; k < e; ++k) {
            const int2 a = c.adj[k];
            const int o = a.x & 0x7fffffff;
            incidence(a, S[o], Qn[o]);
        }
    }
    if (kind >= 3 && cnt > 0 && w != 0.0f) {
        const float md = kind == 3 ? fmaxf(5.0e-4f, fminf(fmaxf(pv.thick * 0.20f, 1.0e-3f), 0.005f))
                                   : fminf(fmaxf(pv.thick * 0.015f, 2.5e-4f), 9.0e-4f);
        F3 d = mul(sum, 1.0f / static_cast<float>(cnt));
        if (!finite3(d)) {
            atomicAdd(&c.counts[kCountHardFallbacks], 1ull);
        } else {
            d = limit_delta(d, md);
            if (finite3(d) && len3(d) > 0.0f) {
                p = p + d;
                q = q + mul(d, kind == 3 ? 0.16f : 0.24f);
            }
        }
    }
    D[v] = with_w(p, w);
    c.prev[v] = with_w(q, 0.0f);
}

// One colour of a cap kind, one thread per edge, in place (race free for a
// valid colouring).  A segment is {cloth, first edge, end edge}.
__global__ __launch_bounds__(kTile) void k_edge_color(const ClothView* __restrict__ views, const int4* __restrict__ segs, int nsegs,
                                                      int parity, int kind) {
    const int4 seg = segs[tile_of_block(nsegs)];
    const int k = seg.y + static_cast<int>(threadIdx.x);
    if (k >= seg.z) return;
    const ClothView& c = views[seg.x];
    const PcgView& pv = c.pcg;
    const float dt = c.dt_sub;
    float4* P = cur(c, parity);
    const int2 ed = pv.edges[k];
    const int i = ed.x, j = ed.y;
    if (i >= pv.nv || j >= pv.nv || i == j) return;
    const float4 pi4 = P[i], pj4 = P[j];
    const F3 qi = f3(c.prev[i]), qj = f3(c.prev[j]);
    const float rest = pv.edge_rest[k];
    EdgeOut oi, oj;
    const bool ai = edge_kind(kind, f3(pi4), f3(pj4), qi, qj, pi4.w, pj4.w, rest, pv.thick, pv.strength_cfg, dt, &oi) && oi.apply;
    const bool aj = edge_kind(kind, f3(pj4), f3(pi4), qj, qi, pj4.w, pi4.w, rest, pv.thick, pv.strength_cfg, dt, &oj) && oj.apply;
    if (ai) { P[i] = with_w(f3(pi4) + oi.d, pi4.w); c.prev[i] = with_w(qi + oi.q, 0.0f); }
    if (aj) { P[j] = with_w(f3(pj4) + oj.d, pj4.w); c.prev[j] = with_w(qj + oj.q, 0.0f); }
}

// Zeroes the scalars, guard and max-delta of every cloth of the group.
__global__ void k_pcg_reset(const ClothView* __restrict__ views, const int* __restrict__ members, int n) {
    const int m = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    if (m >= n) return;
    const PcgView& pv = views[members[m]].pcg;
    *pv.sc = PcgScalars{0.0, 0.0, 0.0, 0.0, 0.0f, 0, 0, 0};
    *pv.guard = 0ull;
    *pv.max_delta = 0.0f;
}

}  // namespace

// ------------------------------------------------------------------ host
PcgPath::~PcgPath() {
    for (auto& a : allocations_) device_free(a.first, a.second);
}

void* PcgPath::alloc_bytes(size_t bytes) {
    const size_t n = std::max<size_t>(bytes, 256);
    void* p = device_alloc(n, "PCG allocation");
    if (p) allocations_.emplace_back(p, n);
    return p;
}

bool PcgPath::prepare(const SsblXpbdConfig& cfg, const std::vector<int>& edges, const std::vector<float>& edge_rest,
                      const std::vector<int>& colors, int nv, int nadj, int ntiles, int ell_width, int ell_pitch,
                      int ell_overflow, unsigned long long* counts) {
    ready_ = false;
    if (!cfg.stretch_optimization_enabled) return true;
    if (nadj <= 0) return fail("PCG stretch optimization found no valid CSR edges");
    const int ne = cfg.edge_count;
    nnz_ = nadj;
    ntiles_ = ntiles;
    PcgView& v = pv_;
    v = PcgView{};
    v.nv = nv;
    v.ne = ne;
    v.strength_cfg = cfg.stretch_optimization_strength;
    v.ntiles = ntiles;
    v.strength = std::min(std::max(cfg.stretch_optimization_strength * 2.5f, 0.0f), 1.0f);
    v.thick = cfg.cloth_thickness;
    v.hard = (cfg.stretch_optimization_strength >= 0.95f && std::isfinite(cfg.cloth_thickness) && cfg.cloth_thickness > 1.0e-7f) ? 1 : 0;
    v.wb_limit = v.hard ? std::max(2.5e-4f, std::min(cfg.cloth_thickness * 0.035f, 9.0e-4f)) : 0.25f;   // A41:1091-1105
    v.counts = counts;
    v.ell_off = (ell_overflow == 0 && ell_width > 0) ? 1 : 0;
    off_blocks_ = v.ell_off ? static_cast<size_t>(ell_width) * static_cast<size_t>(ell_pitch) : static_cast<size_t>(nadj);
    auto f4 = [&](float4** p) { *p = static_cast<float4*>(alloc_bytes(sizeof(float4) * nv)); return *p != nullptr; };
    bool ok = f4(&v.rhs) && f4(&v.sol) && f4(&v.res) && f4(&v.z) && f4(&v.dir) && f4(&v.adir);
    v.diag = static_cast<float*>(alloc_bytes(sizeof(float) * 6 * static_cast<size_t>(nv)));
    v.pinv = static_cast<float*>(alloc_bytes(sizeof(float) * 6 * static_cast<size_t>(nv)));
    v.off = static_cast<float*>(alloc_bytes(sizeof(float) * 6 * off_blocks_));
    v.partials = static_cast<unsigned long long*>(alloc_bytes(sizeof(unsigned long long) * 2 * static_cast<size_t>(ntiles)));
    v.sc = static_cast<PcgScalars*>(alloc_bytes(sizeof(PcgScalars)));
    v.guard = static_cast<unsigned long long*>(alloc_bytes(sizeof(unsigned long long)));
    v.max_delta = static_cast<float*>(alloc_bytes(sizeof(float)));
    v.tile_max = static_cast<float*>(alloc_bytes(sizeof(float) * static_cast<size_t>(ntiles)));
    int2* d_edges = static_cast<int2*>(alloc_bytes(sizeof(int2) * std::max(ne, 1)));
    float* d_rest = static_cast<float*>(alloc_bytes(sizeof(float) * std::max(ne, 1)));
    if (!ok || !v.diag || !v.pinv || !v.off || !v.partials || !v.sc || !v.guard || !v.max_delta || !v.tile_max || !d_edges || !d_rest)
        return false;
    std::vector<int2> he(ne);
    for (int e = 0; e < ne; ++e) he[e] = make_int2(edges[2 * e], edges[2 * e + 1]);
    if (ne > 0 && (!hip_ok(hipMemcpy(d_edges, he.data(), sizeof(int2) * ne, hipMemcpyHostToDevice), "PCG edge upload") ||
                   !hip_ok(hipMemcpy(d_rest, edge_rest.data(), sizeof(float) * ne, hipMemcpyHostToDevice), "PCG edge upload")))
        return false;
    v.edges = d_edges;
    v.edge_rest = d_rest;
    if (!hip_ok(hipMemset(v.sc, 0, sizeof(PcgScalars)), "PCG reset") || !hip_ok(hipMemset(v.guard, 0, 8), "PCG reset") ||
        !hip_ok(hipMemset(v.max_delta, 0, 4), "PCG reset"))
        return false;
    // colour segments for the coloured cap passes; a colouring that puts a
    // vertex twice in one colour cannot use the race-free per-colour form.
    // Without colours the caps are one Jacobi gather over every edge.
    colored_ = !colors.empty();
    segs_host_.clear();
    seg_first_.clear();
    seg_count_.clear();
    if (colored_) {
        std::vector<int> stamp(nv, -1);
        const int ncol = static_cast<int>(colors.size()) - 1;
        for (int c = 0; c < ncol; ++c) {
            const int lo = std::max(colors[c], 0), hi = std::min(colors[c + 1], ne);
            for (int e = lo; e < hi; ++e) {
                const int i = edges[2 * e], j = edges[2 * e + 1];
                if (i >= nv || j >= nv) continue;
                if (stamp[i] == c || stamp[j] == c)
                    return fail("the hard-stretch caps need a valid edge colouring (no vertex twice in one colour)");
                stamp[i] = c; stamp[j] = c;
            }
        }
        for (int c = 0; c < ncol; ++c) {
            const int lo = std::max(colors[c], 0), hi = std::min(colors[c + 1], ne);
            seg_first_.push_back(static_cast<int>(segs_host_.size()));
            for (int e = lo; e < hi; e += kTile) segs_host_.push_back(make_int4(0, e, std::min(hi, e + kTile), 0));
            seg_count_.push_back(static_cast<int>(segs_host_.size()) - seg_first_.back());
        }
    }
    if (!colored_ && v.strength_cfg >= 0.95f && ne > 0) {
        v.qsnap = static_cast<float4*>(alloc_bytes(sizeof(float4) * nv));
        if (!v.qsnap) return false;
    }
    own_ = pv_;
    ready_ = true;
    return true;
}

// Scratch layout: six float4 vectors and two 3x3 symmetric blocks per vertex,
// one block per directed edge, two 64-bit words per tile, the scalars; each
// part starts on a 256-byte boundary.
namespace {
size_t align256(size_t n) { return (n + 255) & ~size_t(255); }
}  // namespace

size_t PcgPath::scratch_bytes() const {
    const size_t nv = static_cast<size_t>(own_.nv);
    return 6 * align256(sizeof(float4) * nv) + 2 * align256(sizeof(float) * 6 * nv) +
           align256(sizeof(float) * 6 * off_blocks_) + align256(sizeof(unsigned long long) * 2 * static_cast<size_t>(ntiles_)) +
           align256(sizeof(PcgScalars)) + 2 * align256(sizeof(unsigned long long)) + align256(sizeof(float) * static_cast<size_t>(ntiles_)) +
           (own_.qsnap ? align256(sizeof(float4) * nv) : 0);
}

bool PcgPath::use_scratch(void* base, size_t bytes) {
    if (!ready_) return true;
    const size_t need = scratch_bytes();
    if (need > bytes)
        return fail("the hard-stretch solve of this cloth needs " + std::to_string(need) + " bytes of scratch; the static scratch holds " +
                    std::to_string(bytes) + " bytes");
    char* p = static_cast<char*>(base);
    auto take = [&](size_t n) { char* q = p; p += align256(n); return q; };
    PcgView v = own_;
    const size_t nv = static_cast<size_t>(own_.nv);
    v.rhs = reinterpret_cast<float4*>(take(sizeof(float4) * nv));
    v.sol = reinterpret_cast<float4*>(take(sizeof(float4) * nv));
    v.res = reinterpret_cast<float4*>(take(sizeof(float4) * nv));
    v.z = reinterpret_cast<float4*>(take(sizeof(float4) * nv));
    v.dir = reinterpret_cast<float4*>(take(sizeof(float4) * nv));
    v.adir = reinterpret_cast<float4*>(take(sizeof(float4) * nv));
    v.diag = reinterpret_cast<float*>(take(sizeof(float) * 6 * nv));
    v.pinv = reinterpret_cast<float*>(take(sizeof(float) * 6 * nv));
    v.off = reinterpret_cast<float*>(take(sizeof(float) * 6 * off_blocks_));
    v.partials = reinterpret_cast<unsigned long long*>(take(sizeof(unsigned long long) * 2 * static_cast<size_t>(ntiles_)));
    v.sc = reinterpret_cast<PcgScalars*>(take(sizeof(PcgScalars)));
    v.guard = reinterpret_cast<unsigned long long*>(take(sizeof(unsigned long long)));
    v.max_delta = reinterpret_cast<float*>(take(sizeof(unsigned long long)));
    v.tile_max = reinterpret_cast<float*>(take(sizeof(float) * static_cast<size_t>(ntiles_)));
    if (own_.qsnap) v.qsnap = reinterpret_cast<float4*>(take(sizeof(float4) * nv));
    pv_ = v;
    return true;
}

// ------------------------------------------------------------- groups
// Grow-only: a group rebuilt for new members reuses its arrays when they fit.
template <class T> bool PcgGroup::Buf<T>::upload(const std::vector<T>& h, const char* what) {
    if (h.empty()) return true;
    if (h.size() > n) {
        device_free(p, sizeof(T) * n);
        p = static_cast<T*>(device_alloc(sizeof(T) * h.size(), what));
        n = p ? h.size() : 0;
        if (!p) return false;
    }
    return hip_ok(hipMemcpy(p, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice), what);
}

bool PcgGroup::build(const std::vector<PcgMember>& members, int cadence) {
    cadence_ = std::max(cadence, 1);
    std::vector<Tile> tiles, hard_c, hard_u;
    std::vector<int> idx;
    std::vector<int4> segs;
    size_t colors = 0;
    for (const PcgMember& m : members) {
        idx.push_back(m.index);
        const bool polishes = m.path->polishes();
        for (int t = 0; t * kTile < m.nv; ++t) {
            const Tile tile{m.index, t * kTile, std::min(m.nv, (t + 1) * kTile), 0};
            tiles.push_back(tile);
            if (polishes) (m.path->colored() ? hard_c : hard_u).push_back(tile);
        }
        if (polishes && m.path->colored()) colors = std::max(colors, m.path->cap_first().size());
    }
    // colour c of the caps: colour c's segments of every coloured cloth that polishes
    cap_first_.clear();
    cap_count_.clear();
    for (size_t c = 0; c < colors; ++c) {
        cap_first_.push_back(static_cast<int>(segs.size()));
        for (const PcgMember& m : members) {
            if (!m.path->polishes() || !m.path->colored() || c >= m.path->cap_first().size()) continue;
            for (int k = 0; k < m.path->cap_count()[c]; ++k) {
                int4 seg = m.path->cap_segments()[m.path->cap_first()[c] + k];
                seg.x = m.index;
                segs.push_back(seg);
            }
        }
        cap_count_.push_back(static_cast<int>(segs.size()) - cap_first_.back());
    }
    ntiles_ = static_cast<int>(tiles.size());
    nmembers_ = static_cast<int>(idx.size());
    colored_.colored = true;
    colored_.ntiles = static_cast<int>(hard_c.size());
    plain_.colored = false;
    plain_.ntiles = static_cast<int>(hard_u.size());
    return tiles_.upload(tiles, "PCG tile table") && colored_.tiles.upload(hard_c, "PCG tile table") &&
           plain_.tiles.upload(hard_u, "PCG tile table") && members_.upload(idx, "PCG member table") &&
           cap_segs_.upload(segs, "PCG cap segments");
}

void PcgGroup::reset(hipStream_t st, const ClothView* views) const {
    if (nmembers_ == 0) return;
    TimedLaunch t("pcg", st);
    launch_k(k_pcg_reset, dim3((nmembers_ + 63) / 64), dim3(64), 0, st, views, members_.p, nmembers_);
}

void PcgGroup::solve(hipStream_t st, const ClothView* views, int parity) const {
    if (ntiles_ == 0) return;
    TimedLaunch t("pcg", st);
    const dim3 grid(ntiles_), block(kTile), rgrid(nmembers_), rblock(kReduceBlock);
    launch_k(k_pcg_build, grid, block, 0, st, views, tiles_.p, ntiles_, parity);
    launch_k(k_pcg_reduce, rgrid, rblock, 0, st, views, members_.p, 0, 0);
    for (int it = 0; it < kPcgIterations; ++it) {
        const int more = it + 1 < kPcgIterations ? 1 : 0;
        launch_k(k_pcg_ad, grid, block, 0, st, views, tiles_.p, ntiles_, parity);
        launch_k(k_pcg_reduce, rgrid, rblock, 0, st, views, members_.p, 1, more);
        launch_k(k_pcg_update, grid, block, 0, st, views, tiles_.p, ntiles_, parity);
        launch_k(k_pcg_reduce, rgrid, rblock, 0, st, views, members_.p, 2, more);
        if (more) launch_k(k_pcg_dir, grid, block, 0, st, views, tiles_.p, ntiles_, parity);
    }
}

// A gather over one polishing set's tiles (src -> dst); returns the new parity.
int PcgGroup::gather(hipStream_t st, const ClothView* views, const Polishing& set, int parity, int kind) const {
    TimedLaunch t("edge_gather", st);
    launch_k(k_edge_gather, dim3(set.ntiles), dim3(kTile), 0, st, views, set.tiles.p, set.ntiles, parity, kind);
    return parity ^ 1;
}

// Caps of one kind, `passes` times: for coloured cloths each colour in place,
// forward and, when asked, backward (A41:2249-2362, 2496-2708); for the
// others one Jacobi gather over every edge per pass.  Returns the new parity.
int PcgGroup::caps(hipStream_t st, const ClothView* views, const Polishing& set, int parity, int kind, int passes,
                   bool reverse_too) const {
    if (!set.colored) {
        for (int p = 0; p < passes; ++p) {
            if (kind <= 2) {
                TimedLaunch t("edge_gather", st);
                launch_k(k_edge_gather, dim3(set.ntiles), dim3(kTile), 0, st, views, set.tiles.p, set.ntiles, parity, -2);
            }
            parity = gather(st, views, set, parity, kind);
        }
        return parity;
    }
    TimedLaunch t("edge_color", st);
    const int ncol = static_cast<int>(cap_first_.size());
    for (int p = 0; p < passes; ++p) {
        for (int c = 0; c < ncol; ++c)
            if (cap_count_[c] > 0)
                launch_k(k_edge_color, dim3(cap_count_[c]), dim3(kTile), 0, st, views, cap_segs_.p + cap_first_[c], cap_count_[c],
                         parity, kind);
        if (reverse_too)
            for (int c = ncol - 1; c >= 0; --c)
                if (cap_count_[c] > 0)
                    launch_k(k_edge_color, dim3(cap_count_[c]), dim3(kTile), 0, st, views, cap_segs_.p + cap_first_[c], cap_count_[c],
                             parity, kind);
    }
    return parity;
}

// A41:7823-7857 polish sequence, then the extreme caps (A41:7974-7995).  The
// gathers flip the polishing cloths' buffers; an odd count is copied back.
void PcgGroup::polish(hipStream_t st, const ClothView* views, int parity) const {
    for (const Polishing* set : {&colored_, &plain_}) {
        if (set->ntiles == 0) continue;
        int p = parity;
        for (int k = 0; k < 2; ++k) {
            p = gather(st, views, *set, p, 3);
            p = caps(st, views, *set, p, 0, 2, true);
            for (int t = 0; t < 2; ++t) {
                p = gather(st, views, *set, p, 4);
                p = caps(st, views, *set, p, 1, 3, false);
            }
        }
        p = caps(st, views, *set, p, 2, 3, false);
        if (p != parity) gather(st, views, *set, p, -1);
    }
}

// A41:7997-8027
void PcgGroup::final_caps(hipStream_t st, const ClothView* views, int parity) const {
    for (const Polishing* set : {&colored_, &plain_}) {
        if (set->ntiles == 0) continue;
        int p = caps(st, views, *set, parity, 0, 2, true);
        p = caps(st, views, *set, p, 2, 3, false);
        p = caps(st, views, *set, p, 1, 3, false);
        if (p != parity) gather(st, views, *set, p, -1);
    }
}

bool PcgPath::fetch(SsblXpbdDiagnostics* d) const {
    if (!ready_) return true;
    PcgScalars s{};
    unsigned long long guard = 0;
    float maxd = 0.0f;
    if (!hip_ok(hipMemcpy(&s, pv_.sc, sizeof(s), hipMemcpyDeviceToHost), "fetch PCG") ||
        !hip_ok(hipMemcpy(&guard, pv_.guard, 8, hipMemcpyDeviceToHost), "fetch PCG") ||
        !hip_ok(hipMemcpy(&maxd, pv_.max_delta, 4, hipMemcpyDeviceToHost), "fetch PCG"))
        return false;
    d->abi41_pcg_csr_nnz = nnz_;
    d->abi41_pcg_texture_ready = 1;
    d->abi41_pcg_iterations = 8LL * s.runs;
    if (s.runs > 0) {   // A41:7586-7624
        double init = s.initial;
        if ((!std::isfinite(init) || init <= 0.0) && std::isfinite(s.rz) && s.rz >= 0.0) init = s.rz;
        if (std::isfinite(init) && init >= 0.0) {
            d->abi41_pcg_initial_residual = static_cast<float>(std::sqrt(std::max(init, 0.0)));
            d->abi41_pcg_final_residual = d->abi41_pcg_initial_residual;
        }
        const double fin = s.status == 2 ? s.rz : s.rznext;
        if ((s.status == 0 || s.status == 2) && std::isfinite(fin) && fin >= 0.0)
            d->abi41_pcg_final_residual = static_cast<float>(std::sqrt(std::max(fin, 0.0)));
        else if (s.status != 0 && s.status != 2)
            d->abi41_pcg_guarded += 1;
        d->abi41_pcg_guarded += static_cast<long long>(guard);
        d->abi41_pcg_max_delta = maxd;
    }
    return true;
}

}  // namespace ssbl
