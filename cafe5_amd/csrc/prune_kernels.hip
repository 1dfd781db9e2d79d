// prune_kernels.hip -- the plain FP64 pruning kernel (the drop-in for
// inference_prune / compute_node_probability, src/core.cpp:134-145,
// src/probability.cpp:175-330), recompute-list packing, and the per-family
// and per-trial reductions of base_model.cpp:77-95 and gamma_core.cpp:143-225.
// The batched MFMA pruning kernels are in prune_mfma.hip.
//
// Pruning kernels are persistent: a fixed grid pulls TileDesc work items from
// a device counter until the device-resident tile count is reached, so every
// wave exits once the list is drained.
#include "engine.h"

namespace cafe {

// Next work item of a one-wave workgroup; a plain read first, so that once the
// list is drained the remaining workgroups leave without queueing on the
// counter.
__device__ __forceinline__ int fetch_tile(const PruneParams& p, int n_tiles)
{
    int t = n_tiles;
    if (threadIdx.x == 0) {
        const int seen = *(volatile int*)p.tile_counter;
        if (seen < n_tiles) t = atomicAdd(p.tile_counter, 1);
    }
    return __shfl(t, 0);
}

// ---------------------------------------------------------------------------
// Plain FP64 pruning: the drop-in.  One wave per family at a time, lanes over
// rows s; every matrix-vector product runs over x = 0..B in ascending order
// with separately rounded multiply and add, as matrix::multiply
// (matrix_cache.cpp:50-56) does, and the node product starts from 1
// (probability.cpp:208-217, 222-231).  Partials live in per-wave scratch.
// NQ: rows per lane (row s = lane + 64 q).
// ---------------------------------------------------------------------------
template <int NQ>
__global__ __launch_bounds__(64) void k_prune_plain(PruneParams p)
{
    const int lane = threadIdx.x;
    extern __shared__ double s_plain_slots[];
    // partial vectors live in LDS when the launch provides it (n_slots x ldv
    // doubles), else in this workgroup's region of the global scratch
    double* slots = p.plain_lds ? s_plain_slots : (double*)p.scratch + (long long)blockIdx.x * p.scratch_per_wg;
    const int ldv = p.ldv;
    const int Np = p.Np;
    const int n_tiles = *p.n_tiles_dev;

    for (int tile = fetch_tile(p, n_tiles); tile < n_tiles; tile = fetch_tile(p, n_tiles)) {
        const TileDesc td = p.tiles[tile];
        const BucketDev bk = p.buckets[td.bucket];
        const int* __restrict__ nslot = p.node_slot + (long long)td.point * p.n_nodes;
        const int B = bk.B;
        const bool packed = td.nf < 0;
        const int nf = packed ? -td.nf : td.nf;
        for (int fi = 0; fi < nf; ++fi) {
            const int fs = packed ? p.fam_list[td.fam0 + fi] : bk.fam_begin + td.fam0 + fi;
            double best = -__builtin_huge_val();
            double xmax = 0.0;
            for (int op = 0; op < p.n_ops; ++op) {
                const OpDev o = p.ops[op];
                const bool is_root = (op == p.n_ops - 1);
                const int len = is_root ? bk.R : bk.L;
                const int shift = is_root ? 1 : 0;
                double acc[NQ];
#pragma unroll
                for (int q = 0; q < NQ; ++q) acc[q] = 1.0;
                for (int k = 0; k < o.n_child; ++k) {
                    const ChildDev ch = p.kids[o.child_begin + k];
                    const double* __restrict__ M = p.mt64 + (long long)nslot[ch.node] * p.slot_elems + shift;
                    double x[NQ];
#pragma unroll
                    for (int q = 0; q < NQ; ++q) x[q] = 0.0;
                    if (ch.leaf_col >= 0) {
                        // column count of the leaf matrix (error model folded in:
                        // the same sum in the same order as the leaf vector's
                        // matrix-vector product)
                        const int cnt = p.counts_T[(long long)ch.leaf_col * p.F + fs];
                        const double* __restrict__ Ml = p.leaf64 + (long long)nslot[ch.node] * p.slot_elems + shift;
#pragma unroll
                        for (int q = 0; q < NQ; ++q) {
                            const int s = lane + 64 * q;
                            if (s < len) x[q] = Ml[(long long)cnt * Np + s];
                        }
                    } else {
                        // x[s] += M[s][c] v[c] for c = 0..B in ascending order; the
                        // loads of U columns are issued before their products, so
                        // U column reads are in flight instead of one
                        constexpr int U = 32 / NQ < 8 ? 32 / NQ : 8;
                        const double* V = slots + (long long)ch.in_slot * ldv;
                        for (int c0 = 0; c0 <= B; c0 += U) {
                            double vv[U], cv[U][NQ];
#pragma unroll
                            for (int u = 0; u < U; ++u) {
                                const int c = c0 + u <= B ? c0 + u : B;
                                vv[u] = V[c];
                                const double* col = M + (long long)c * Np;
#pragma unroll
                                for (int q = 0; q < NQ; ++q) {
                                    const int s = lane + 64 * q;
                                    cv[u][q] = s < len ? col[s] : 0.0;
                                }
                            }
#pragma unroll
                            for (int u = 0; u < U; ++u) {
                                if (c0 + u > B) break;
#pragma unroll
                                for (int q = 0; q < NQ; ++q) {
                                    const int s = lane + 64 * q;
                                    if (s < len) x[q] = __dadd_rn(x[q], __dmul_rn(cv[u][q], vv[u]));
                                }
                            }
                        }
                    }
#pragma unroll
                    for (int q = 0; q < NQ; ++q) acc[q] = __dmul_rn(acc[q], x[q]);
                }
                if (!is_root) {
                    __syncthreads();   // all lanes finished reading the child slots
                    double* out = slots + (long long)o.out_slot * ldv;
#pragma unroll
                    for (int q = 0; q < NQ; ++q) {
                        const int s = lane + 64 * q;
                        if (s < len) out[s] = acc[q];
                    }
                    __syncthreads();
                } else {
                    const double* pr = p.prior64 + bk.prior_off;
                    int bj = -1;
#pragma unroll
                    for (int q = 0; q < NQ; ++q) {
                        const int j = lane + 64 * q;
                        if (j >= len) continue;
                        if (p.root_out)
                            p.root_out[(long long)p.perm[fs] * p.root_stride + j] = acc[q];
                        // base_model.cpp:86 / gamma_core.cpp:157
                        const double key = p.model == 0 ? log(acc[q]) + log(pr[j]) : __dmul_rn(acc[q], pr[j]);
                        if (bj < 0 || best < key) {
                            best = key;
                            bj = j;
                        }
                        xmax = xmax > acc[q] ? xmax : acc[q];
                    }
                    for (int d = 1; d < 64; d <<= 1) {
                        const double ob = __shfl_xor(best, d);
                        const double ox = __shfl_xor(xmax, d);
                        best = ob > best ? ob : best;
                        xmax = ox > xmax ? ox : xmax;
                    }
                }
            }
            if (lane == 0) {
                p.out_val[(long long)td.point * p.F + p.perm[fs]] = best;
                p.out_flag[(long long)td.point * p.F + p.perm[fs]] = xmax == 0.0 ? kFlagZeroSum : 0;
            }
            __syncthreads();
        }
    }
}

// Packs the families of one segment chunk (up to kSegChunk families of a
// bucket at one point) whose flag has `mask` set into dense tiles of `tile`
// families (16 for MFMA passes, 1 for PLAIN64): their bucket-sorted indices
// go to the chunk's own region of fam_list (point * F + fam_begin + first, so
// no two chunks collide), and the tiles are appended to the recompute list of
// the chunk's size class.
__global__ __launch_bounds__(256) void k_pack_flags(const SegDesc* __restrict__ segs,
                                                    const BucketDev* __restrict__ buckets,
                                                    const int* __restrict__ perm, const int* __restrict__ flags,
                                                    int F, int mask, int* __restrict__ fam_list,
                                                    TileDesc* __restrict__ out_tiles,
                                                    const int* __restrict__ class_off, int* __restrict__ class_count,
                                                    int* __restrict__ fam_count, int tile)
{
    const SegDesc sg = segs[blockIdx.x];
    const BucketDev bk = buckets[sg.bucket];
    const long long base = (long long)sg.point * F + bk.fam_begin + sg.first;
    const int count = bk.count - sg.first < kSegChunk ? bk.count - sg.first : kSegChunk;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    __shared__ int wave_tot[4];
    __shared__ int s_total;
    if (tid == 0) s_total = 0;
    __syncthreads();
    for (int i0 = 0; i0 < count; i0 += 256) {
        const int i = i0 + tid;
        const bool hit = i < count && (flags[(long long)sg.point * F + perm[bk.fam_begin + sg.first + i]] & mask) != 0;
        const unsigned long long ball = __ballot(hit);
        const int before = __popcll(ball & ((1ull << lane) - 1ull));
        if (lane == 0) wave_tot[wave] = __popcll(ball);
        __syncthreads();
        int off = s_total;
        for (int w = 0; w < wave; ++w) off += wave_tot[w];
        if (hit) fam_list[base + off + before] = bk.fam_begin + sg.first + i;
        __syncthreads();
        if (tid == 0) s_total += wave_tot[0] + wave_tot[1] + wave_tot[2] + wave_tot[3];
        __syncthreads();
    }
    const int n = s_total;
    if (n == 0) return;
    const int nt = (n + tile - 1) / tile;
    __shared__ int s_first;
    if (tid == 0) {
        s_first = atomicAdd(&class_count[sg.cls], nt);
        atomicAdd(fam_count, n);
    }
    __syncthreads();
    for (int k = tid; k < nt; k += 256) {
        const int left = n - tile * k;
        out_tiles[class_off[sg.cls] + s_first + k] =
            TileDesc{sg.bucket, sg.point, (int)(base + tile * k), -(left < tile ? left : tile)};
    }
}

// Per (trial, family) model reduction: base copies lnL; gamma combines the K
// categories as gamma_core.cpp:190-208 (cat_k = (max_j L_j prior_j) p_k,
// family = sum_k cat_k, posterior_k = cat_k p_k / sum_k cat_k p_k).
__global__ void k_finalize(int model, int T, int K, int F, const double* __restrict__ val,
                           const int* __restrict__ flag, const double* __restrict__ cat_probs,
                           double* __restrict__ lnl, double* __restrict__ cat, double* __restrict__ post,
                           int* __restrict__ fail)
{
    const long long n = (long long)T * F;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        const int t = (int)(i / F);
        const int f = (int)(i % F);
        if (model == 0) {
            lnl[i] = val[i];
            fail[i] = 0;
            continue;
        }
        bool zero = false;
        for (int k = 0; k < K; ++k) zero |= (flag[((long long)t * K + k) * F + f] & kFlagZeroSum) != 0;
        fail[i] = zero ? 1 : 0;
        double fam = 0.0, denom = 0.0;
        for (int k = 0; k < K; ++k) {
            const double c = zero ? 0.0 : __dmul_rn(val[((long long)t * K + k) * F + f], cat_probs[k]);
            cat[i * K + k] = c;
            fam = __dadd_rn(fam, c);
            denom = __dadd_rn(denom, __dmul_rn(c, cat_probs[k]));
        }
        for (int k = 0; k < K; ++k)
            post[i * K + k] = __ddiv_rn(__dmul_rn(cat[i * K + k], cat_probs[k]), denom);
        lnl[i] = zero ? __builtin_nan("") : log(fam);
    }
}

// total[t] = -sum_f lnl[t][f] with compensated (Neumaier) summation over
// each thread's strided share and a two-sum tree; +inf when any gamma family
// failed.
__device__ __forceinline__ void two_sum(double a, double b, double& s, double& e)
{
    s = __dadd_rn(a, b);
    const double bb = __dsub_rn(s, a);
    e = __dadd_rn(__dsub_rn(a, __dsub_rn(s, bb)), __dsub_rn(b, bb));
}

constexpr int kTotalsThreads = 1024;

__global__ __launch_bounds__(kTotalsThreads) void k_totals(int F, const double* __restrict__ lnl,
                                                const int* __restrict__ fail, double* __restrict__ total)
{
    const int t = blockIdx.x;
    const double* x = lnl + (long long)t * F;
    const int* fl = fail + (long long)t * F;
    const int tid = threadIdx.x;
    // non-finite values are tallied apart (two-sum of an infinity gives NaN):
    // bit 1 gamma failure, 2 NaN, 4 -inf, 8 +inf.  Thread tid takes families
    // tid, tid + kTotalsThreads, ... (the block's loads coalesce).
    double s = 0.0, c = 0.0;
    int bad = 0;
#pragma unroll 4
    for (int i = tid; i < F; i += kTotalsThreads) {
        const double v = x[i];
        if (fl[i]) bad |= 1;
        if (v != v) {
            bad |= 2;
        } else if (v == -__builtin_huge_val()) {
            bad |= 4;
        } else if (v == __builtin_huge_val()) {
            bad |= 8;
        } else {
            double ns, e;
            two_sum(s, v, ns, e);
            s = ns;
            c = __dadd_rn(c, e);
        }
    }
    __shared__ double ss[kTotalsThreads], sc[kTotalsThreads];
    __shared__ int sb[kTotalsThreads];
    ss[tid] = s;
    sc[tid] = c;
    sb[tid] = bad;
    __syncthreads();
    for (int w = kTotalsThreads / 2; w > 0; w >>= 1) {
        if (tid < w) {
            double ns, e;
            two_sum(ss[tid], ss[tid + w], ns, e);
            ss[tid] = ns;
            sc[tid] = __dadd_rn(__dadd_rn(sc[tid], sc[tid + w]), e);
            sb[tid] |= sb[tid + w];
        }
        __syncthreads();
    }
    if (tid == 0) {
        const int b = sb[0];
        const double inf = __builtin_huge_val();
        double r;
        if (b & 1) r = inf;                             // gamma_core.cpp:216-225
        else if ((b & 2) || ((b & 4) && (b & 8))) r = __builtin_nan("");
        else if (b & 4) r = inf;                        // -sum with a -inf term
        else if (b & 8) r = -inf;
        else r = -__dadd_rn(ss[0], sc[0]);
        total[t] = r;
    }
}

// ---------------------------------------------------------------------------
// host-side launchers
// ---------------------------------------------------------------------------
// PLAIN64 launch: nq rows per lane (row s = lane + 64 q), one wave per
// workgroup; partial vectors in dynamic LDS when they fit in 48 KB.
hipError_t launch_prune_plain(int nq, int grid, const PruneParams& p, hipStream_t st)
{
    if (grid < 1) grid = 1;
    const size_t lds = sizeof(double) * (size_t)p.n_slots * p.ldv;
    PruneParams q = p;
    q.plain_lds = lds <= 48 * 1024 ? 1 : 0;
    const size_t shm = q.plain_lds ? lds : 0;
    if (nq <= 1) hipLaunchKernelGGL(k_prune_plain<1>, dim3(grid), dim3(64), shm, st, q);
    else if (nq <= 2) hipLaunchKernelGGL(k_prune_plain<2>, dim3(grid), dim3(64), shm, st, q);
    else if (nq <= 4) hipLaunchKernelGGL(k_prune_plain<4>, dim3(grid), dim3(64), shm, st, q);
    else if (nq <= 8) hipLaunchKernelGGL(k_prune_plain<8>, dim3(grid), dim3(64), shm, st, q);
    else if (nq <= 16) hipLaunchKernelGGL(k_prune_plain<16>, dim3(grid), dim3(64), shm, st, q);
    else if (nq <= 32) hipLaunchKernelGGL(k_prune_plain<32>, dim3(grid), dim3(64), shm, st, q);
    else return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t launch_pack(const SegDesc* segs, int n_segs, const BucketDev* buckets, const int* perm,
                       const int* flags, int F, int mask, int* fam_list, TileDesc* out_tiles,
                       const int* class_off, int* class_count, int* fam_count, int tile, hipStream_t st)
{
    if (n_segs <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_pack_flags, dim3(n_segs), dim3(256), 0, st, segs, buckets, perm, flags, F, mask, fam_list,
                       out_tiles, class_off, class_count, fam_count, tile);
    return hipGetLastError();
}

hipError_t launch_finalize(int model, int T, int K, int F, const double* val, const int* flag,
                           const double* cat_probs, double* lnl, double* cat, double* post, int* fail,
                           double* total, hipStream_t st)
{
    long long n = (long long)T * F;
    long long blocks = (n + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(k_finalize, dim3((unsigned)blocks), dim3(256), 0, st, model, T, K, F, val, flag,
                       cat_probs, lnl, cat, post, fail);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_totals, dim3(T), dim3(kTotalsThreads), 0, st, F, lnl, fail, total);
    return hipGetLastError();
}

}  // namespace cafe
