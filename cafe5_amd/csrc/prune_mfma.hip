// prune_mfma.hip -- batched Felsenstein pruning on the matrix cores
// (replaces inference_prune / compute_node_probability, src/core.cpp:134-145,
// src/probability.cpp:175-330, and the root reductions of base_model.cpp:77-95
// and gamma_core.cpp:143-165, for many families at once).
//
// A workgroup of W waves carries 16 families (the columns of every
// contraction) of one size bucket at one parameter point through the whole
// tree, one internal node after another (Sethi-Ullman order, root last):
//     v_u[s] = prod_c  sum_x M_c[s][x] v_c[x]            s = 0..B
//     root:  v[j] = prod_c sum_x M_c[j+1][x] v_c[x]       j = 0..R-1
// A leaf child contributes column M_c[.][count] (or the error-model mix of
// neighbouring columns), which is what the reference's matrix-vector product
// with a one-hot leaf vector evaluates to.
//
// Fragment layout.  Lane l = 16 g + f of a wave holds, for 16-row tile t of a
// vector, the four sizes 16 t + 4 g + r (r = 0..3) of family f as one 4-vector.
// This is the D layout of v_mfma_f32_16x16x4f32, and it is also the B operand
// layout of the next contraction when the contraction index of MFMA j of
// k-tile kt in lane group g is x = 16 kt + 4 g + j: a node's output is the next
// node's input without any data movement.  v_mfma_f64_16x16x4f64 returns D row
// g + 4 r in register r; the FP64 kernel feeds A row m from size
// 4 (m % 4) + m / 4, which lands size 4 g + r in that register as well.
//
// Persistent kernels: a fixed grid pulls tiles from a device counter until the
// device-resident tile count is reached; every wave of a workgroup sees the
// same tile index (broadcast through LDS), so all waves leave together.
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

__device__ __forceinline__ const float* pool_of(const PruneParams& p, float*) { return p.mt32; }
__device__ __forceinline__ const double* pool_of(const PruneParams& p, double*) { return p.mt64; }
__device__ __forceinline__ const float* err_of(const PruneParams& p, float*) { return p.err32; }
__device__ __forceinline__ const double* err_of(const PruneParams& p, double*) { return p.err64; }

// Size that A row m of a 16-row tile carries (see the fragment layout above).
__device__ __forceinline__ int a_row(float*, int m) { return m; }
__device__ __forceinline__ int a_row(double*, int m) { return 4 * (m & 3) + (m >> 2); }

// Matrix element (s, c): the FP32 pool is row-major, the FP64 pool c-major.
// Offsets within one matrix fit in 32 bits (Np <= 1040), so addresses are a
// uniform base plus a 32-bit lane offset.
__device__ __forceinline__ float mat(const float* M, int Np, int s, int c) { return M[s * Np + c]; }
__device__ __forceinline__ double mat(const double* M, int Np, int s, int c) { return M[c * Np + s]; }

// A operand of k-tile kt for a lane whose A row is size s: the four entries
// (s, 16 kt + 4 g + j), j = 0..3.  FP32: one 16-byte load from the row.
__device__ __forceinline__ int a_off(float*, int Np, int s, int g) { return s * Np + 4 * g; }
__device__ __forceinline__ floatx4 load_a(const float* M, int off, int kt, int)
{
    return *(const floatx4*)(M + off + kTile * kt);
}
__device__ __forceinline__ int a_off(double*, int Np, int s, int g) { return 4 * g * Np + s; }
__device__ __forceinline__ doublex4 load_a(const double* M, int off, int kt, int Np)
{
    const int o = off + kTile * kt * Np;
    return doublex4{M[o], M[o + Np], M[o + 2 * Np], M[o + 3 * Np]};
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
    double key;   // x * prior (scaled path: in units of 2^e_sum)
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

__device__ __forceinline__ void root_merge(RootBest& b, double key, double x, int j, double xmax)
{
    if (j >= 0) {
        const double xm = b.xmax > xmax ? b.xmax : xmax;
        root_take(b, key, x, j);
        b.xmax = xm;
    }
}

__device__ __forceinline__ RootBest root_combine(RootBest b)
{
    for (int d = 16; d <= 32; d <<= 1) {
        const double ok = __shfl_xor(b.key, d);
        const double ox = __shfl_xor(b.x, d);
        const int oj = __shfl_xor(b.j, d);
        const double om = __shfl_xor(b.xmax, d);
        root_merge(b, ok, ox, oj, om);
    }
    return b;
}

// One k-tile of MFMAs: x[i] += A_i(kt) B(kt) for the wave's NTW output tiles.
// SCALED: the stored child fragment is renormalised on the fly by 2^bsh.
template <typename T, int NTW, bool SCALED>
__device__ __forceinline__ void mma_step(typename Vec4<T>::type (&x)[NTW], const typename Vec4<T>::type (&a)[NTW],
                                         const typename Vec4<T>::type& b, int bsh)
{
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        T bj = b[j];
        if constexpr (SCALED) bj = ldexpf(bj, bsh);
#pragma unroll
        for (int i = 0; i < NTW; ++i) x[i] = mfma16(a[i][j], bj, x[i]);
    }
}

// x[i] = sum_x M(srow_i, x) v(x) over the child's nts k-tiles, two k-tiles
// per iteration in ping-pong buffers: the loads of one buffer are in flight
// while the other is multiplied.  The body has no branches (a k-tile past the
// end re-reads the last one with its B fragment zeroed, which adds exactly 0),
// so each wait covers only the older buffer's loads.
template <typename T, int NTW, bool SCALED>
__device__ __forceinline__ void child_gemm(typename Vec4<T>::type (&x)[NTW], const T* __restrict__ M, int Np,
                                           const int (&srow)[NTW], int g, const typename Vec4<T>::type* bsrc,
                                           int nts, int bsh)
{
    typedef typename Vec4<T>::type V4;
    int ao[NTW];
#pragma unroll
    for (int i = 0; i < NTW; ++i) ao[i] = a_off((T*)nullptr, Np, srow[i], g);
    V4 a0[NTW], a1[NTW], b0, b1;
#pragma unroll
    for (int i = 0; i < NTW; ++i) a0[i] = load_a(M, ao[i], 0, Np);
    b0 = bsrc[0];
    for (int kt = 0; kt < nts; kt += 2) {
        const bool has1 = kt + 1 < nts;
        const int k1 = has1 ? kt + 1 : kt;
#pragma unroll
        for (int i = 0; i < NTW; ++i) a1[i] = load_a(M, ao[i], k1, Np);
        b1 = bsrc[k1 * 64];
        mma_step<T, NTW, SCALED>(x, a0, b0, bsh);
        const int k2 = kt + 2 < nts ? kt + 2 : nts - 1;
#pragma unroll
        for (int i = 0; i < NTW; ++i) a0[i] = load_a(M, ao[i], k2, Np);
        b0 = bsrc[k2 * 64];
        if (!has1) b1 = V4{0, 0, 0, 0};
        mma_step<T, NTW, SCALED>(x, a1, b1, bsh);
    }
}

// ---------------------------------------------------------------------------
// The pruning kernel.  Wave wv computes output tiles t0 + wv + W i, i < NTW
// (FP64 makes several passes when a vector has more than W NTW tiles; tiles
// past the vector's end are computed on a clamped, in-bounds tile and
// dropped).  Partial vectors are kept in fragment layout, in LDS (LDSP) or in
// the workgroup's region of the global scratch; one barrier per node makes a
// node's output visible to every wave.
//
// T = float, SCALED: FP32 matrices are stored times 2^kMatExp; a node's raw
// product is stored as computed, and its per-family maximum m (over all waves)
// renormalises it when it is read: v' = raw 2^(a - e), e = exponent of m, so
// max v' is just below 2^a (a = vec_exp).  The exponent is carried as an
// integer, and a per-family bound on what FP32 flushing could have dropped
// decides at the root whether the family is kept or recomputed in FP64.
// T = double: unscaled, so values (and underflow) follow the reference.
// ---------------------------------------------------------------------------
template <typename T, int NTW, int W, bool SCALED, bool LDSP>
__global__ __launch_bounds__(64 * W) void k_prune_mfma(PruneParams p)
{
    typedef typename Vec4<T>::type V4;
    const int lane = threadIdx.x & 63;
    const int wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int fl = lane & 15;
    const int g = lane >> 4;
    const int my_arow = a_row((T*)nullptr, fl);
    const T* __restrict__ pool = pool_of(p, (T*)nullptr);
    const T* __restrict__ err = err_of(p, (T*)nullptr);
    const int Np = p.Np;
    const int n_slots = p.n_slots;
    const int slot_tiles = p.slot_tiles;
    const int a_exp = p.vec_exp;

    // dynamic LDS: scaled-path bookkeeping per slot, then LDS partials
    extern __shared__ __align__(32) unsigned char s_dyn[];
    double* m_cpre = (double*)s_dyn;                  // [slot][16] flush bound before renormalising
    float* m_wmax = (float*)(m_cpre + n_slots * 16);  // [slot][W][16] per-wave maxima
    int* m_esum = (int*)(m_wmax + n_slots * W * 16);  // [slot][16] exponent of the raw product
    int* m_l2s = m_esum + n_slots * 16;               // [slot] log2 of the nominal scale
    V4* frag = LDSP ? (V4*)(s_dyn + mfma_meta_bytes(n_slots, W, SCALED))
                    : (V4*)p.scratch + (long long)blockIdx.x * n_slots * slot_tiles * 64;
    __shared__ double s_rkey[W][16], s_rx[W][16], s_rxmax[W][16];
    __shared__ int s_rj[W][16];
    __shared__ int s_tile;

    const int n_tiles = __builtin_amdgcn_readfirstlane(*p.n_tiles_dev);
    for (;;) {
        if (threadIdx.x == 0) {
            // a plain read first: once the list is drained, workgroups leave
            // without queueing on the counter
            int t = *(volatile int*)p.tile_counter;
            s_tile = t < n_tiles ? atomicAdd(p.tile_counter, 1) : n_tiles;
        }
        __syncthreads();
        const int tile = __builtin_amdgcn_readfirstlane(s_tile);
        __syncthreads();   // s_tile is rewritten only after every wave has read it
        if (tile >= n_tiles) break;
        const TileDesc td = p.tiles[tile];
        const BucketDev bk = p.buckets[td.bucket];
        const int nf = td.nf < 0 ? -td.nf : td.nf;
        const bool valid = fl < nf;
        const int fi = valid ? fl : 0;
        const int fs = td.nf < 0 ? p.fam_list[td.fam0 + fi] : bk.fam_begin + td.fam0 + fi;
        const int dst = p.perm[fs];   // the caller's index of this lane's family
        const int* __restrict__ nslot = p.node_slot + (long long)td.point * p.n_nodes;
        const int nts = bk.Lp / kTile;   // k-tiles of every child vector

        for (int op = 0; op < p.n_ops; ++op) {
            const OpDev o = p.ops[op];
            const bool is_root = op == p.n_ops - 1;
            const int rows = is_root ? bk.Rp / kTile : nts;
            const int limit = is_root ? bk.R : bk.L;
            const int shift = is_root ? 1 : 0;

            // SCALED bookkeeping.  A renormalised child v' holds v_c = v' 2^E_c
            // with max v' < 2^a; its factor X' = M' v' (M' = 2^kMatExp M) holds
            // X_c = X' 2^(E_c - kMatExp) and is at most 2^(kMatExp + a) (leaf
            // factors: 2^kMatExp).
            //   e_sum  sum_c (E_c - kMatExp): true product = raw 2^e_sum
            //   log2S  log2 of the product's nominal scale prod_c 2^(kMatExp + a_c)
            //   cert   sum_c of the factors' relative flush errors (w.r.t. their
            //          nominal scale): inherited bound + what flushing of M'
            //          entries, MFMA products and stored v' entries can drop
            int e_sum = 0, log2S = 0;
            double cert = 0.0;
            RootBest rb{-__builtin_huge_val(), 0.0, -1, 0.0};
            T wmax = 0;

            for (int t0 = 0; t0 < rows; t0 += W * NTW) {
                int tt[NTW], srow[NTW];
                bool live[NTW];
#pragma unroll
                for (int i = 0; i < NTW; ++i) {
                    const int t = t0 + wv + W * i;
                    live[i] = t < rows;
                    tt[i] = live[i] ? t : rows - 1;
                    srow[i] = kTile * tt[i] + my_arow + shift;
                }
                V4 acc[NTW];
#pragma unroll
                for (int i = 0; i < NTW; ++i) acc[i] = V4{1, 1, 1, 1};
                for (int k = 0; k < o.n_child; ++k) {
                    const ChildDev ch = p.kids[o.child_begin + k];
                    const T* __restrict__ M = pool + (long long)nslot[ch.node] * p.slot_elems;
                    V4 x[NTW];
#pragma unroll
                    for (int i = 0; i < NTW; ++i) x[i] = V4{0, 0, 0, 0};
                    if (ch.leaf_col >= 0) {
                        if (SCALED && t0 == 0) {
                            e_sum -= kMatExp;
                            log2S += kMatExp;
                            cert += 3.0 * kTau32 * kTwoPowMinusMatExp;   // up to 3 mixed columns
                        }
                        const int cnt = p.counts_T[(long long)ch.leaf_col * p.F + fs];
                        if (p.n_dev == 0) {
#pragma unroll
                            for (int i = 0; i < NTW; ++i)
#pragma unroll
                                for (int r = 0; r < 4; ++r)
                                    x[i][r] = mat(M, Np, kTile * tt[i] + 4 * g + r + shift, cnt);
                        } else {
                            // error model: columns cnt-h .. cnt+h weighted by
                            // get_probs(cnt); columns below 0 get weight 0
                            const int h = (p.n_dev - 1) / 2;
                            for (int d = 0; d < p.n_dev; ++d) {
                                const int cc = cnt - h + d;
                                const T w = cc >= 0 ? err[(long long)cnt * p.n_dev + d] : T(0);
                                const int c2 = cc >= 0 ? cc : 0;
#pragma unroll
                                for (int i = 0; i < NTW; ++i)
#pragma unroll
                                    for (int r = 0; r < 4; ++r)
                                        x[i][r] += w * mat(M, Np, kTile * tt[i] + 4 * g + r + shift, c2);
                            }
                        }
                    } else {
                        int bsh = 0;
                        if (SCALED) {
                            float m = 0.f;
                            for (int w2 = 0; w2 < W; ++w2) m = fmaxf(m, m_wmax[(ch.in_slot * W + w2) * 16 + fl]);
                            int e = 0;
                            const double mant = m > 0.f ? frexp((double)m, &e) : 1.0;
                            bsh = a_exp - e;
                            if (t0 == 0) {
                                e_sum += m_esum[ch.in_slot * 16 + fl] + e - a_exp - kMatExp;
                                log2S += kMatExp + a_exp;
                                const double c_in = m > 0.f ? ldexp(m_cpre[ch.in_slot * 16 + fl] * 1.01 / mant,
                                                                    m_l2s[ch.in_slot] - e) +
                                                                  ldexp(kTau32 / mant, -e)
                                                            : 1e300;
                                cert += c_in + kTau32 * (bk.Lp * kTwoPowMinusMatExp * (1.0 + ldexp(1.0, -a_exp)) +
                                                         ldexp(1.0, -a_exp));
                            }
                        }
                        const V4* bsrc = frag + (long long)ch.in_slot * slot_tiles * 64 + lane;
                        child_gemm<T, NTW, SCALED>(x, M, Np, srow, g, bsrc, nts, bsh);
                    }
#pragma unroll
                    for (int i = 0; i < NTW; ++i) acc[i] *= x[i];
                }

                if (!is_root) {
#pragma unroll
                    for (int i = 0; i < NTW; ++i) {
                        if (!live[i]) continue;
                        V4 v = acc[i];
#pragma unroll
                        for (int r = 0; r < 4; ++r)
                            if (kTile * tt[i] + 4 * g + r >= limit) v[r] = 0;   // rows past the bound
                        frag[((long long)o.out_slot * slot_tiles + tt[i]) * 64 + lane] = v;
                        if (SCALED)
#pragma unroll
                            for (int r = 0; r < 4; ++r) wmax = wmax > v[r] ? wmax : v[r];
                    }
                } else {
                    // prior values first (unconditional, clamped), then the reduction
                    const double* __restrict__ pr = p.prior64 + bk.prior_off;
                    double pv[NTW][4];
#pragma unroll
                    for (int i = 0; i < NTW; ++i)
#pragma unroll
                        for (int r = 0; r < 4; ++r) {
                            const int j = kTile * tt[i] + 4 * g + r;
                            pv[i][r] = pr[j < limit ? j : limit - 1];
                        }
#pragma unroll
                    for (int i = 0; i < NTW; ++i) {
                        if (!live[i]) continue;
#pragma unroll
                        for (int r = 0; r < 4; ++r) {
                            const int j = kTile * tt[i] + 4 * g + r;
                            if (j >= limit) continue;
                            const double xv = (double)acc[i][r];
                            if (p.root_out && valid)
                                p.root_out[(long long)dst * p.root_stride + j] = SCALED ? ldexp(xv, e_sum) : xv;
                            root_take(rb, xv * pv[i][r], xv, j);
                        }
                    }
                }
            }

            if (!is_root) {
                if (SCALED) {
                    float m = (float)wmax;
                    m = fmaxf(m, __shfl_xor(m, 16));
                    m = fmaxf(m, __shfl_xor(m, 32));
                    if (g == 0) m_wmax[(o.out_slot * W + wv) * 16 + fl] = m;
                    if (wv == 0 && g == 0) {
                        m_esum[o.out_slot * 16 + fl] = e_sum;
                        m_cpre[o.out_slot * 16 + fl] = cert;
                    }
                    if (wv == 0 && lane == 0) m_l2s[o.out_slot] = log2S;
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
                    if (wv == 0)
                        for (int w2 = 1; w2 < W; ++w2)
                            root_merge(rb, s_rkey[w2][fl], s_rx[w2][fl], s_rj[w2][fl], s_rxmax[w2][fl]);
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
                            if (!(log(rb.key) + log_scale >= log(kFp64Floor))) flag |= kFlagRecompute64;
                        }
                    } else {
                        // base: log(x_j prior_j) at the best j (base_model.cpp:86
                        // takes log x_j + log prior_j); gamma: x_j prior_j
                        val = p.model == 0 ? log(rb.key) : rb.key;
                        if (rb.xmax == 0.0) flag |= kFlagZeroSum;
                        if (rb.xmax < kFp64Floor || rb.key < kFp64Floor) flag |= kFlagRecomputePlain;
                    }
                    if (valid && g == 0) {
                        p.out_val[(long long)td.point * p.F + dst] = val;
                        p.out_flag[(long long)td.point * p.F + dst] = flag;
                    }
                }
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Launcher: one kernel per (tiles per wave, waves per workgroup, partials in
// LDS) shape; the host picks the shape per size class.
// ---------------------------------------------------------------------------
template <typename T, bool SCALED>
static hipError_t launch_shape(const MfmaShape& s, int grid, const PruneParams& p, hipStream_t st)
{
#define CAFE_MFMA_CASE(NTW, W, LDSP)                                                                    \
    if (s.ntw == NTW && s.w == W && s.lds == LDSP) {                                                    \
        hipLaunchKernelGGL((k_prune_mfma<T, NTW, W, SCALED, LDSP>), dim3(grid), dim3(64 * W), s.shm, st, p); \
        return hipGetLastError();                                                                       \
    }
#define CAFE_MFMA_BOTH(NTW, W) CAFE_MFMA_CASE(NTW, W, true) CAFE_MFMA_CASE(NTW, W, false)
    CAFE_MFMA_BOTH(1, 1)
    CAFE_MFMA_BOTH(2, 1)
    CAFE_MFMA_BOTH(3, 1)
    CAFE_MFMA_BOTH(4, 1)
    CAFE_MFMA_BOTH(3, 2)
    CAFE_MFMA_BOTH(4, 2)
    CAFE_MFMA_BOTH(3, 4)
    CAFE_MFMA_BOTH(4, 4)
    CAFE_MFMA_CASE(3, 8, false)
    CAFE_MFMA_CASE(4, 8, false)
    if constexpr (SCALED) {   // FP64 covers large vectors in several passes of (4, 8)
        CAFE_MFMA_CASE(3, 16, false)
        CAFE_MFMA_CASE(4, 16, false)
    }
#undef CAFE_MFMA_BOTH
#undef CAFE_MFMA_CASE
    return hipErrorInvalidValue;
}

hipError_t launch_prune_mfma(bool fp64, const MfmaShape& s, int grid, const PruneParams& p, hipStream_t st)
{
    if (grid < 1) grid = 1;
    return fp64 ? launch_shape<double, false>(s, grid, p, st) : launch_shape<float, true>(s, grid, p, st);
}

}  // namespace cafe
