// prune_kernels.hip -- Felsenstein pruning of many gene families on MI355X
// (replaces inference_prune / compute_node_probability, src/core.cpp:134-145,
// src/probability.cpp:175-330, and the root reductions of base_model.cpp:77-95
// and gamma_core.cpp:143-165).
//
// All pruning kernels are persistent: a fixed grid of single-wave workgroups
// pulls TileDesc work items from a device counter until the device-resident
// tile count is reached, so every wave exits once the list is drained.
//
// Per internal node u with children c (program order from the host, root last):
//     v_u[s] = prod_c  sum_x M_c[s][x] v_c[x]            s = 0..B
//     root:  v[j] = prod_c sum_x M_c[j+1][x] v_c[x]       j = 0..R-1
// A leaf child contributes the column M_c[.][count] (or the error-model mix of
// neighbouring columns); that is exactly what the reference's matrix-vector
// product with a one-hot leaf vector evaluates to.
#include "engine.h"

namespace cafe {

typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef double doublex4 __attribute__((ext_vector_type(4)));

template <typename T> struct Vec4;
template <> struct Vec4<float> { typedef floatx4 type; };
template <> struct Vec4<double> { typedef doublex4 type; };

__device__ __forceinline__ floatx4 mfma16(float a, float b, floatx4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ doublex4 mfma16(double a, double b, doublex4 c)
{
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// Output row held by lane group g (= lane / 16), register r, of 16-row tile t.
// v_mfma_f32_16x16x4f32 puts D rows 4g..4g+3 in one lane; v_mfma_f64_16x16x4f64
// interleaves them (rows g, g+4, g+8, g+12), as measured on gfx950.
__device__ __forceinline__ int drow(float*, int t, int g, int r) { return 16 * t + 4 * g + r; }
__device__ __forceinline__ int drow(double*, int t, int g, int r) { return 16 * t + g + 4 * r; }

__device__ __forceinline__ const float* pool_of(const PruneParams& p, float*) { return p.mt32; }
__device__ __forceinline__ const double* pool_of(const PruneParams& p, double*) { return p.mt64; }
__device__ __forceinline__ const float* err_of(const PruneParams& p, float*) { return p.err32; }
__device__ __forceinline__ const double* err_of(const PruneParams& p, double*) { return p.err64; }

__device__ __forceinline__ int fetch_tile(const PruneParams& p)
{
    int t = 0;
    if (threadIdx.x == 0) t = atomicAdd(p.tile_counter, 1);
    return __shfl(t, 0);
}

constexpr double kLn2 = 0.69314718055994530942;
// FP32 flush threshold (smallest normal float): the flush bound charges this
// for every quantity FP32 could have flushed to zero.
constexpr double kTau32 = 1.1754943508222875e-38;
constexpr double kTwoPowMinusMatExp = 1.0 / (double)(1 << kMatExp);
// an FP32 likelihood is accepted when the flush bound is below this fraction
constexpr double kCertEta = 1e-9;
// below this, FP64 values may lose bits to subnormals; such families are
// recomputed in the reference's summation order (PLAIN64)
constexpr double kFp64Floor = 1e-290;

// Running root reduction of one lane: best key over its rows.
struct RootBest {
    double key;   // base FP64: log x + log prior; otherwise x * prior
    double x;     // x at the best key
    int j;        // its index (smallest j on ties)
    double xmax;  // max_j x_j
};

__device__ __forceinline__ void root_take(RootBest& b, double key, double x, int j)
{
    if (b.j < 0 || b.key < key || (b.key == key && j < b.j)) {
        b.key = key;
        b.x = x;
        b.j = j;
    }
    b.xmax = b.xmax > x ? b.xmax : x;
}

__device__ __forceinline__ RootBest root_combine(RootBest b)
{
    for (int d = 16; d <= 32; d <<= 1) {
        RootBest o;
        o.key = __shfl_xor(b.key, d);
        o.x = __shfl_xor(b.x, d);
        o.j = __shfl_xor(b.j, d);
        o.xmax = __shfl_xor(b.xmax, d);
        if (o.j >= 0) {
            const double xm = b.xmax > o.xmax ? b.xmax : o.xmax;
            root_take(b, o.key, o.x, o.j);
            b.xmax = xm;
        }
    }
    return b;
}

// ---------------------------------------------------------------------------
// Batched MFMA pruning.  A workgroup of W waves carries 16 families (the
// columns of every contraction) of one bucket at one parameter point through
// the whole tree.  For v_mfma_*_16x16x4: lane l supplies A[s0 + l%16][k] and
// B[k][f = l%16] with k = l/16, and holds four rows of D[.][f = l%16] (rows
// given by drow()).  Within each 16-wide chunk of the contraction index, MFMA
// j maps lane l's k to x = 16 kt + 4 (l/16) + j, so a lane reads four
// consecutive entries of a partial vector of its own family.
//
// The output rows of a node are 16-row tiles; wave w computes tiles
// w, w + W, w + 2W, ... (NTW tiles per pass; the FP64 kernel makes more
// passes when a vector has more tiles).  Partial vectors live in the
// workgroup's scratch region; a barrier after each node makes them visible to
// every wave.  W > 1 shortens the critical path of large buckets, whose
// tiles are few but long.
//
// T = float, SCALED: each node's partial is renormalised per family to a max
// just below 2^vec_exp (FP32 matrices are stored times 2^kMatExp) and the
// exponent is carried as an integer; a per-family bound on what FP32 flushing
// could have dropped decides whether the result is kept.  SCALED needs all of
// a vector's tiles in one pass (W * NTW >= tiles).
// T = double: unscaled, so values (and underflow) follow the reference.
// ---------------------------------------------------------------------------
template <typename T, int NTW, int W, bool SCALED>
__global__ __launch_bounds__(64 * W) void k_prune_mfma(PruneParams p)
{
    typedef typename Vec4<T>::type V4;
    const int lane = threadIdx.x & 63;
    const int wv = threadIdx.x >> 6;
    const int fl = lane & 15;
    const int g = lane >> 4;
    const T* __restrict__ pool = pool_of(p, (T*)nullptr);
    const T* __restrict__ err = err_of(p, (T*)nullptr);
    T* slots = (T*)p.scratch + (long long)blockIdx.x * p.scratch_per_wg;
    const int ldv = p.ldv;
    const int Np = p.Np;
    const int n_tiles = *p.n_tiles_dev;
    __shared__ int s_exp[kMaxSlots][16];       // SCALED: exponent per slot/family
    __shared__ double s_cert[kMaxSlots][16];   // SCALED: flush bound per slot/family
    __shared__ double s_max[W][16];            // per-wave family maxima
    __shared__ double s_rkey[W][16], s_rx[W][16], s_rxmax[W][16];
    __shared__ int s_rj[W][16];
    __shared__ int s_tile;

    for (;;) {
        if (threadIdx.x == 0) s_tile = atomicAdd(p.tile_counter, 1);
        __syncthreads();
        const int tile = s_tile;
        __syncthreads();
        if (tile >= n_tiles) break;
        const TileDesc td = p.tiles[tile];
        const BucketDev bk = p.buckets[td.bucket];
        const bool packed = td.nf < 0;
        const bool valid = fl < (packed ? -td.nf : td.nf);
        const int fs = packed ? p.fam_list[td.fam0 + (valid ? fl : 0)] : bk.fam_begin + td.fam0 + (valid ? fl : 0);
        const int* __restrict__ nslot = p.node_slot + (long long)td.point * p.n_nodes;
        const int nts = bk.Lp / kTile;
        const int a_exp = p.vec_exp;

        for (int op = 0; op < p.n_ops; ++op) {
            const OpDev o = p.ops[op];
            const bool is_root = (op == p.n_ops - 1);
            const int rows = is_root ? bk.Rp / kTile : nts;
            const int limit = is_root ? bk.R : bk.L;
            const int shift = is_root ? 1 : 0;

            // SCALED bookkeeping from the children.  A stored partial v' of
            // child c holds v_c = v' 2^E_c with max v' < 2^a; its factor
            // X' = M' v' (M' = 2^kMatExp M) holds X_c = X' 2^(E_c - kMatExp)
            // and is at most 2^(kMatExp + a) (leaf factors: 2^kMatExp).
            //   e_sum   sum_c (E_c - kMatExp): true product = P' 2^e_sum
            //   log2S   log2 of the product's nominal scale S = prod_c 2^(kMatExp + a_c)
            //   cert    sum_c of the factors' relative flush errors (w.r.t. their
            //           nominal scale): inherited bound + what flushing of M'
            //           entries, MFMA products and stored v' entries can drop
            int e_sum = 0, log2S = 0;
            double cert = 0.0;
            if (SCALED) {
                for (int k = 0; k < o.n_child; ++k) {
                    const ChildDev ch = p.kids[o.child_begin + k];
                    if (ch.leaf_col >= 0) {
                        e_sum -= kMatExp;
                        log2S += kMatExp;
                        cert += 3.0 * kTau32 * kTwoPowMinusMatExp;   // up to 3 mixed columns
                    } else {
                        e_sum += s_exp[ch.in_slot][fl] - kMatExp;
                        log2S += kMatExp + a_exp;
                        cert += s_cert[ch.in_slot][fl] +
                                kTau32 * (bk.Lp * kTwoPowMinusMatExp * (1.0 + ldexp(1.0, -a_exp)) + ldexp(1.0, -a_exp));
                    }
                }
            }
            RootBest rb{-__builtin_huge_val(), 0.0, -1, 0.0};
            double fmax_local = 0.0;
            V4 acc[NTW];

            for (int t0 = 0; t0 < rows; t0 += W * NTW) {
                // this wave's tiles of the pass: t0 + wv + W i
#pragma unroll
                for (int i = 0; i < NTW; ++i) acc[i] = V4{1, 1, 1, 1};
                for (int k = 0; k < o.n_child; ++k) {
                    const ChildDev ch = p.kids[o.child_begin + k];
                    const T* __restrict__ M = pool + (long long)nslot[ch.node] * p.slot_elems;
                    V4 x[NTW];
#pragma unroll
                    for (int i = 0; i < NTW; ++i) x[i] = V4{0, 0, 0, 0};
                    if (ch.leaf_col >= 0) {
                        const int cnt = valid ? p.counts_T[(long long)ch.leaf_col * p.F + fs] : 0;
                        if (p.n_dev == 0) {
                            const T* col = M + (long long)cnt * Np + shift;
#pragma unroll
                            for (int i = 0; i < NTW; ++i) {
                                const int tt = t0 + wv + W * i;
                                if (tt < rows) {
#pragma unroll
                                    for (int r = 0; r < 4; ++r) x[i][r] = col[drow((T*)nullptr, tt, g, r)];
                                }
                            }
                        } else {
                            // error model: columns cnt-h .. cnt+h weighted by
                            // get_probs(cnt), negative columns skipped
                            const int h = (p.n_dev - 1) / 2;
                            for (int d = 0; d < p.n_dev; ++d) {
                                const int cc = cnt - h + d;
                                if (cc < 0) continue;
                                const T w = err[(long long)cnt * p.n_dev + d];
                                const T* col = M + (long long)cc * Np + shift;
#pragma unroll
                                for (int i = 0; i < NTW; ++i) {
                                    const int tt = t0 + wv + W * i;
                                    if (tt < rows) {
#pragma unroll
                                        for (int r = 0; r < 4; ++r) x[i][r] += w * col[drow((T*)nullptr, tt, g, r)];
                                    }
                                }
                            }
                        }
                    } else {
                        const T* V = slots + ((long long)ch.in_slot * 16 + fl) * ldv + 4 * g;
                        for (int kt = 0; kt < nts; ++kt) {
                            const V4 b = *(const V4*)(V + kTile * kt);
#pragma unroll
                            for (int j = 0; j < 4; ++j) {
                                const T* a = M + (long long)(kTile * kt + 4 * g + j) * Np + shift + fl;
#pragma unroll
                                for (int i = 0; i < NTW; ++i) {
                                    const int tt = t0 + wv + W * i;
                                    if (tt < rows) x[i] = mfma16(a[kTile * tt], b[j], x[i]);
                                }
                            }
                        }
                    }
#pragma unroll
                    for (int i = 0; i < NTW; ++i) acc[i] *= x[i];
                }
                // rows past the bound hold real matrix rows: clear them
#pragma unroll
                for (int i = 0; i < NTW; ++i) {
                    const int tt = t0 + wv + W * i;
#pragma unroll
                    for (int r = 0; r < 4; ++r)
                        if (tt >= rows || drow((T*)nullptr, tt, g, r) >= limit) acc[i][r] = 0;
                }
                if (!is_root) {
                    if (SCALED) {
#pragma unroll
                        for (int i = 0; i < NTW; ++i)
#pragma unroll
                            for (int r = 0; r < 4; ++r)
                                fmax_local = fmax_local > (double)acc[i][r] ? fmax_local : (double)acc[i][r];
                    } else {
                        T* out = slots + ((long long)o.out_slot * 16 + fl) * ldv;
#pragma unroll
                        for (int i = 0; i < NTW; ++i) {
                            const int tt = t0 + wv + W * i;
                            if (tt < rows)
#pragma unroll
                                for (int r = 0; r < 4; ++r) out[drow((T*)nullptr, tt, g, r)] = acc[i][r];
                        }
                    }
                } else {
                    const double* pr = p.prior64 + bk.prior_off;
#pragma unroll
                    for (int i = 0; i < NTW; ++i) {
                        const int tt = t0 + wv + W * i;
#pragma unroll
                        for (int r = 0; r < 4; ++r) {
                            const int j = drow((T*)nullptr, tt, g, r);
                            if (tt >= rows || j >= limit) continue;
                            const double xv = (double)acc[i][r];
                            if (p.root_out && valid)
                                p.root_out[(long long)p.perm[fs] * p.root_stride + j] =
                                    SCALED ? ldexp(xv, e_sum) : xv;
                            const double key = (!SCALED && p.model == 0) ? log(xv) + log(pr[j]) : xv * pr[j];
                            root_take(rb, key, xv, j);
                        }
                    }
                }
            }

            if (!is_root) {
                if (SCALED) {
                    // per-family max over all waves, then renormalise to a max in
                    // [2^(a-1), 2^a); the bound is re-expressed relative to the
                    // new scale: it grows by S / m, the factor by which the
                    // product fell short of its nominal scale
                    double m = fmax_local;
                    m = fmax(m, __shfl_xor(m, 16));
                    m = fmax(m, __shfl_xor(m, 32));
                    if (g == 0) s_max[wv][fl] = m;
                    __syncthreads();   // also: every wave has read the children's s_exp/s_cert
                    m = 0.0;
#pragma unroll
                    for (int w2 = 0; w2 < W; ++w2) m = fmax(m, s_max[w2][fl]);
                    int e = 0;
                    const double mant = m > 0 ? frexp(m, &e) : 1.0;
                    T* out = slots + ((long long)o.out_slot * 16 + fl) * ldv;
#pragma unroll
                    for (int i = 0; i < NTW; ++i) {
                        const int tt = wv + W * i;
                        if (tt < rows) {
                            V4 v;
#pragma unroll
                            for (int r = 0; r < 4; ++r) v[r] = ldexpf(acc[i][r], a_exp - e);
                            *(V4*)(out + drow((T*)nullptr, tt, g, 0)) = v;
                        }
                    }
                    if (wv == 0 && g == 0) {
                        s_exp[o.out_slot][fl] = e_sum + e - a_exp;
                        s_cert[o.out_slot][fl] =
                            m > 0 ? ldexp(cert * 1.01 / mant, log2S - e) + ldexp(kTau32 / mant, -e) : 1e300;
                    }
                }
                __syncthreads();   // the node's vector is complete before anyone reads it
            } else {
                rb = root_combine(rb);
                if (W > 1) {
                    if (g == 0) {
                        s_rkey[wv][fl] = rb.key;
                        s_rx[wv][fl] = rb.x;
                        s_rxmax[wv][fl] = rb.xmax;
                        s_rj[wv][fl] = rb.j;
                    }
                    __syncthreads();
                    if (wv == 0) {
                        for (int w2 = 1; w2 < W; ++w2) {
                            const int oj = s_rj[w2][fl];
                            if (oj >= 0) {
                                const double xm = rb.xmax > s_rxmax[w2][fl] ? rb.xmax : s_rxmax[w2][fl];
                                root_take(rb, s_rkey[w2][fl], s_rx[w2][fl], oj);
                                rb.xmax = xm;
                            }
                        }
                    }
                }
                if (wv == 0) {
                    double val;
                    int flag = 0;
                    if (SCALED) {
                        // rb.key = x_j * prior_j in units of 2^e_sum
                        if (!(rb.key > 0)) {
                            flag = kFlagRecompute64;
                            val = 0.0;
                        } else {
                            const double log_scale = (double)e_sum * kLn2;
                            if (p.model == 0)
                                val = log(rb.x) + log(p.prior64[bk.prior_off + rb.j]) + log_scale;
                            else
                                val = ldexp(rb.key, e_sum);
                            // flush bound (absolute, in units of 2^e_sum) against
                            // the likelihood, and FP64 range of the unscaled values
                            const double bound = (ldexp(cert, log2S) + kTau32) * 1.01;
                            if (!(bound <= kCertEta * rb.key)) flag |= kFlagRecompute64;
                            if (!(log(rb.xmax) + log_scale >= log(kFp64Floor))) flag |= kFlagRecompute64;
                        }
                    } else {
                        val = rb.key;
                        if (rb.xmax == 0.0) flag |= kFlagZeroSum;
                        if (rb.xmax < kFp64Floor) flag |= kFlagRecomputePlain;
                    }
                    if (valid && g == 0) {
                        p.out_val[(long long)td.point * p.F + p.perm[fs]] = val;
                        p.out_flag[(long long)td.point * p.F + p.perm[fs]] = flag;
                    }
                }
            }
        }
        __syncthreads();
    }
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

    for (int tile = fetch_tile(p); tile < n_tiles; tile = fetch_tile(p)) {
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
                        const int cnt = p.counts_T[(long long)ch.leaf_col * p.F + fs];
                        if (p.n_dev == 0) {
#pragma unroll
                            for (int q = 0; q < NQ; ++q) {
                                const int s = lane + 64 * q;
                                if (s < len) x[q] = M[(long long)cnt * Np + s];
                            }
                        } else {
                            const int h = (p.n_dev - 1) / 2;
                            for (int i = 0; i < p.n_dev; ++i) {
                                const int cc = cnt - h + i;
                                if (cc < 0) continue;
                                const double w = p.err64[(long long)cnt * p.n_dev + i];
#pragma unroll
                                for (int q = 0; q < NQ; ++q) {
                                    const int s = lane + 64 * q;
                                    if (s < len) x[q] = __dadd_rn(x[q], __dmul_rn(M[(long long)cc * Np + s], w));
                                }
                            }
                        }
                    } else {
                        const double* V = slots + (long long)ch.in_slot * ldv;
                        for (int c = 0; c <= B; ++c) {
                            const double v = V[c];
                            const double* col = M + (long long)c * Np;
#pragma unroll
                            for (int q = 0; q < NQ; ++q) {
                                const int s = lane + 64 * q;
                                if (s < len) x[q] = __dadd_rn(x[q], __dmul_rn(col[s], v));
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

// Packs the families of one (bucket, point) segment whose flag has `mask` set
// into dense tiles of `tile` families (16 for MFMA passes, 1 for PLAIN64): their bucket-sorted indices go to the segment's
// own region of fam_list (point * F + fam_begin, so no two segments collide),
// and the tiles are appended to the recompute list of the segment's size class.
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
    const long long base = (long long)sg.point * F + bk.fam_begin;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    __shared__ int wave_tot[4];
    __shared__ int s_total;
    if (tid == 0) s_total = 0;
    __syncthreads();
    for (int i0 = 0; i0 < bk.count; i0 += 256) {
        const int i = i0 + tid;
        const bool hit = i < bk.count && (flags[(long long)sg.point * F + perm[bk.fam_begin + i]] & mask) != 0;
        const unsigned long long ball = __ballot(hit);
        const int before = __popcll(ball & ((1ull << lane) - 1ull));
        if (lane == 0) wave_tot[wave] = __popcll(ball);
        __syncthreads();
        int off = s_total;
        for (int w = 0; w < wave; ++w) off += wave_tot[w];
        if (hit) fam_list[base + off + before] = bk.fam_begin + i;
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
// contiguous chunks and a two-sum tree; +inf when any gamma family failed.
__device__ __forceinline__ void two_sum(double a, double b, double& s, double& e)
{
    s = __dadd_rn(a, b);
    const double bb = __dsub_rn(s, a);
    e = __dadd_rn(__dsub_rn(a, __dsub_rn(s, bb)), __dsub_rn(b, bb));
}

__global__ __launch_bounds__(256) void k_totals(int F, const double* __restrict__ lnl,
                                                const int* __restrict__ fail, double* __restrict__ total)
{
    const int t = blockIdx.x;
    const double* x = lnl + (long long)t * F;
    const int* fl = fail + (long long)t * F;
    const int tid = threadIdx.x;
    const int chunk = (F + 255) / 256;
    const int lo = tid * chunk, hi = lo + chunk < F ? lo + chunk : F;
    // non-finite values are tallied apart (two-sum of an infinity gives NaN):
    // bit 1 gamma failure, 2 NaN, 4 -inf, 8 +inf
    double s = 0.0, c = 0.0;
    int bad = 0;
    for (int i = lo; i < hi; ++i) {
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
    __shared__ double ss[256], sc[256];
    __shared__ int sb[256];
    ss[tid] = s;
    sc[tid] = c;
    sb[tid] = bad;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
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
// Size class nt = 16-row tiles of the largest vector: (tiles per wave per
// pass, waves per workgroup).  Small vectors run one wave per tile (many
// tiles, throughput); large ones split rows over waves (few, long tiles).
template <typename T, bool SCALED>
static hipError_t launch_mfma_nt(int nt, int max_tiles, int max_waves, const PruneParams& p, hipStream_t st)
{
    // workgroups: one per tile in flight, within the wave budget
#define CAFE_MFMA_CASE(NTW, W)                                                                          \
    {                                                                                                   \
        int grid = max_waves / W < max_tiles ? max_waves / W : max_tiles;                               \
        if (grid < 1) grid = 1;                                                                         \
        hipLaunchKernelGGL((k_prune_mfma<T, NTW, W, SCALED>), dim3(grid), dim3(64 * W), 0, st, p);      \
        return hipGetLastError();                                                                       \
    }
    if (nt <= 2) CAFE_MFMA_CASE(2, 1)
    if (nt <= 4) CAFE_MFMA_CASE(4, 1)
    if (nt <= 8) CAFE_MFMA_CASE(4, 2)
    if (nt <= 16) CAFE_MFMA_CASE(4, 4)
    if (nt <= 32) CAFE_MFMA_CASE(4, 8)
    if (!SCALED) CAFE_MFMA_CASE(4, 8)   // FP64 makes several passes
#undef CAFE_MFMA_CASE
    return hipErrorInvalidValue;
}

// max_tiles: upper bound of the work list; max_waves: waves to keep resident
// (the scratch holds one region per workgroup, at most max_waves of them).
hipError_t launch_prune(int path, int nt, int max_tiles, int max_waves, const PruneParams& p, hipStream_t st)
{
    int grid = max_tiles < max_waves ? max_tiles : max_waves;
    if (grid < 1) grid = 1;
    if (path == 0) {
        const int nq = nt;   // rows per lane for the plain kernel
        const size_t lds = sizeof(double) * (size_t)p.n_slots * p.ldv;
        PruneParams q = p;
        q.plain_lds = lds <= 48 * 1024 ? 1 : 0;
        const size_t shm = q.plain_lds ? lds : 0;
        if (nq <= 1) hipLaunchKernelGGL(k_prune_plain<1>, dim3(grid), dim3(64), shm, st, q);
        else if (nq <= 2) hipLaunchKernelGGL(k_prune_plain<2>, dim3(grid), dim3(64), shm, st, q);
        else if (nq <= 4) hipLaunchKernelGGL(k_prune_plain<4>, dim3(grid), dim3(64), shm, st, q);
        else if (nq <= 8) hipLaunchKernelGGL(k_prune_plain<8>, dim3(grid), dim3(64), shm, st, q);
        else if (nq <= 16) hipLaunchKernelGGL(k_prune_plain<16>, dim3(grid), dim3(64), shm, st, q);
        else return hipErrorInvalidValue;
        return hipGetLastError();
    }
    if (path == 1) return launch_mfma_nt<double, false>(nt, max_tiles, max_waves, p, st);
    return launch_mfma_nt<float, true>(nt, max_tiles, max_waves, p, st);
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
    hipLaunchKernelGGL(k_totals, dim3(T), dim3(256), 0, st, F, lnl, fail, total);
    return hipGetLastError();
}

}  // namespace cafe
