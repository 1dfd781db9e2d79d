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
// Persistent kernels: a fixed grid walks the class's tile list (its length is
// device-resident, so recompute lists built on the device need no host round
// trip); every wave of a workgroup works on the same tile.
#include "prune_common.h"

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
__device__ __forceinline__ const float* leaf_of(const PruneParams& p, float*) { return p.leaf32; }
__device__ __forceinline__ const double* leaf_of(const PruneParams& p, double*) { return p.leaf64; }

// Waits for this wave's LDS traffic and meets the workgroup at a barrier,
// leaving global loads in flight (the next step's prefetched operands).
__device__ __forceinline__ void lds_barrier()
{
    asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
}

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

// ds_swizzle pattern: lane l reads lane l ^ 16 (within each 32-lane half).
constexpr int kSwizzleXor16 = 0x401F;

// max(v, v of lane l ^ 32): v_permlane32_swap exchanges the upper half of its
// first operand with the lower half of its second.
__device__ __forceinline__ float xor32_max(float v)
{
    const unsigned u = __builtin_bit_cast(unsigned, v);
    const auto sw = __builtin_amdgcn_permlane32_swap(u, u, false, false);
    return fmaxf(__builtin_bit_cast(float, (unsigned)sw[0]), __builtin_bit_cast(float, (unsigned)sw[1]));
}

// Makes the wave wait for a loaded value without using it (an empty asm that
// reads each element), so that no load of it is left in flight.
template <typename V>
__device__ __forceinline__ void await_loaded(const V& v)
{
#pragma unroll
    for (int j = 0; j < 4; ++j) asm volatile("" ::"v"(v[j]));
}

// Operand prefetch depth: FP32 prefetches k-tiles 0 and 1 of an internal
// step (two 16-byte loads per output tile), FP64 k-tile 0 (four 8-byte loads).
template <typename T> struct Prefetch { static constexpr int PK = sizeof(T) == 4 ? 2 : 1; };

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

// One k-tile of MFMAs: x[f][i] += A_i(kt) B_f(kt) for the wave's NTW output
// tiles and the tile's FB blocks of 16 families (each A operand feeds FB
// MFMAs).  SCALED: the stored child fragment of block f is renormalised on the
// fly by 2^bsh[f].
template <typename T, int NTW, bool SCALED, int FB>
__device__ __forceinline__ void mma_step(typename Vec4<T>::type (&x)[FB][NTW], const typename Vec4<T>::type (&a)[NTW],
                                         const typename Vec4<T>::type (&b)[FB], const int (&bsh)[FB])
{
#pragma unroll
    for (int j = 0; j < 4; ++j) {
#pragma unroll
        for (int f = 0; f < FB; ++f) {
            T bj = b[f][j];
            if constexpr (SCALED) bj = ldexpf(bj, bsh[f]);
#pragma unroll
            for (int i = 0; i < NTW; ++i) x[f][i] = mfma16(a[i][j], bj, x[f][i]);
        }
    }
}

// x[i] = sum_x M(srow_i, x) v(x) over the child's nts k-tiles.  The first
// PK k-tiles' A operands arrive prefetched (pf, loaded during the previous
// step); the remaining ones (vectors of more than PK 16-row tiles) stream
// through a ring of D buffers, D k-tiles' loads in flight while one is
// multiplied (the large vectors read their matrices from HBM).  Loads are
// unconditional: a k-tile past the end re-reads the last one with its B
// fragment zeroed, which adds exactly 0, so each wait covers only the oldest
// buffer.  A one-wave kernel holds whole vectors (nts <= NTW), so its ring
// only needs the NTW - PK k-tiles past the prefetched ones (none for NTW <= PK).
template <typename T, int NTW, int W, int FB> struct Ring {
    // two k-tiles in flight; a deeper ring costs registers (the 4-tile FP32
    // shapes would drop to one wave per SIMD) without shortening their step
    // time, which other waves' work covers.  The FP64 recompute shapes with
    // one or two tiles per wave run a few workgroups whose step time is the
    // k-loop's memory latency: four in flight.  W = 16 leaves 128 VGPRs per
    // wave: two with one output tile of one family block per wave (the FP64
    // recompute of the largest vectors is such a chain), else one.
    static constexpr int deep = W >= 16 ? (NTW == 1 && FB == 1 ? 2 : 1) : (sizeof(T) == 8 && NTW <= 2 ? 4 : 2);
    static constexpr int tail = NTW > Prefetch<T>::PK ? NTW - Prefetch<T>::PK : 0;
    static constexpr int D = W == 1 ? tail : deep;
};

// Child fragments of k-tile kt, family block f: stored at (kt FB + f) 64 + lane
// from the slot's base (both blocks of a k-tile side by side).
template <typename T, int NTW, int D, bool SCALED, int FB>
__device__ __forceinline__ void child_gemm(typename Vec4<T>::type (&x)[FB][NTW], const T* __restrict__ M, int Np,
                                           const int (&ao)[NTW],
                                           const typename Vec4<T>::type (&pf)[Prefetch<T>::PK][NTW],
                                           const typename Vec4<T>::type* bsrc, int nts, const int (&bsh)[FB])
{
    typedef typename Vec4<T>::type V4;
    constexpr int PK = Prefetch<T>::PK;
    {
        V4 b0[FB];
#pragma unroll
        for (int f = 0; f < FB; ++f) b0[f] = bsrc[f * 64];
        mma_step<T, NTW, SCALED, FB>(x, pf[0], b0, bsh);
        if (PK > 1 && nts > 1) {
            V4 b1[FB];
#pragma unroll
            for (int f = 0; f < FB; ++f) b1[f] = bsrc[(FB + f) * 64];
            mma_step<T, NTW, SCALED, FB>(x, pf[PK - 1], b1, bsh);
        } else if (PK > 1) {
            // unused second k-tile: still wait for it (see the leaf step)
#pragma unroll
            for (int i = 0; i < NTW; ++i) await_loaded(pf[PK - 1][i]);
        }
    }
    if constexpr (D == 0) return;
    if (nts <= PK) return;
    V4 a[D > 0 ? D : 1][NTW], bb[D > 0 ? D : 1][FB];
#pragma unroll
    for (int d = 0; d < D; ++d) {
        const int k = PK + d < nts ? PK + d : nts - 1;
#pragma unroll
        for (int i = 0; i < NTW; ++i) a[d][i] = load_a(M, ao[i], k, Np);
#pragma unroll
        for (int f = 0; f < FB; ++f) bb[d][f] = PK + d < nts ? bsrc[(k * FB + f) * 64] : V4{0, 0, 0, 0};
    }
    for (int kt = PK; kt < nts; kt += D) {
#pragma unroll
        for (int d = 0; d < D; ++d) {
            mma_step<T, NTW, SCALED, FB>(x, a[d], bb[d], bsh);
            const int kn = kt + D + d;   // the k-tile this buffer holds next
            const int k = kn < nts ? kn : nts - 1;
#pragma unroll
            for (int i = 0; i < NTW; ++i) a[d][i] = load_a(M, ao[i], k, Np);
#pragma unroll
            for (int f = 0; f < FB; ++f) bb[d][f] = kn < nts ? bsrc[(k * FB + f) * 64] : V4{0, 0, 0, 0};
        }
    }
}

// ---------------------------------------------------------------------------
// The pruning kernel.  The evaluation program is walked as a list of child
// steps (StepDev, op order); while step s is multiplied, the loads of step
// s+1 -- its leaf column, or the first PK k-tiles of its A operand -- are in
// flight in the other half of a ping-pong register buffer (both kinds issue
// the same loads from per-lane offsets computed once per tile, so the waits
// count the younger loads exactly).  Wave wv computes NTW output tiles (W NTW
// covers the largest vector of the size class; tiles past a vector's end are
// computed on a clamped, in-bounds tile and dropped): FP64 tiles wv + W i;
// FP32 the contiguous run [wv n / W, (wv + 1) n / W) of the n row tiles, so
// that an interior step multiplies only the k-tiles inside the union of its
// rows' bands (the FP32 pool's nonzero k-tiles per row tile, recorded by the
// build): the A blocks it skips are exactly 0, so each sum is unchanged.
// Partial vectors are kept in fragment layout, in LDS (LDSP) or in the
// workgroup's region of the global scratch; the leaf counts of the tile's
// families are staged in LDS.
//
// T = float, SCALED: FP32 matrices are stored times 2^kMatExp; a node's raw
// product is stored as computed, and its per-family maximum m renormalises it
// when it is read: v' = raw 2^(a - e), e = exponent of m, so max v' is just
// below 2^a (a = vec_exp).  The exponent is carried as an integer, and a
// per-family bound on what FP32 flushing could have dropped decides at the
// root whether the family is kept or recomputed in FP64.  A single-wave
// workgroup finishes a node's exponent, shift and bound when it stores the
// node; with several waves the reader combines the per-wave maxima.
// T = double: unscaled, so values (and underflow) follow the reference.
// ---------------------------------------------------------------------------
// Single-wave kernels with small per-wave tiles ask for 4 waves per SIMD (the
// register budget that still holds both prefetch buffers without spilling).
template <int NTW, int W, int BYTES> struct WavesHint {
    static constexpr int v = (W == 1 && NTW * BYTES <= 8) ? 4 : 1;
};

// FB = 2: a tile carries 32 families, two blocks of 16 (block f holds
// families 16 f .. 16 f + 15 in lanes fl of every lane group): each A operand
// (matrix fragment) feeds both blocks' MFMAs, so a tile reads its matrices
// once for twice the families, and a k-tile's loads in flight cover twice the
// matrix-core work.  The per-family bookkeeping is kept per block.
template <typename T, int NTW, int W, bool SCALED, bool LDSP, int FB>
__global__ __launch_bounds__(64 * W) __attribute__((amdgpu_waves_per_eu(WavesHint<NTW, W, (int)sizeof(T) * FB>::v)))
void k_prune_mfma(PruneParams p)
{
    typedef typename Vec4<T>::type V4;
    constexpr int PK = Prefetch<T>::PK;
    constexpr bool kF64 = sizeof(T) == 8;
    constexpr int FW = 16 * FB;   // families per tile
    static_assert(FB == 1 || PK >= FB, "a leaf step loads one column per family block into the prefetch buffers");
    const int lane = threadIdx.x & 63;
    const int wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int fl = lane & 15;
    const int g = lane >> 4;
    const int my_arow = a_row((T*)nullptr, fl);
    const T* __restrict__ pool = pool_of(p, (T*)nullptr);
    const T* __restrict__ lpool = leaf_of(p, (T*)nullptr);
    const int Np = p.Np;
    const int n_slots = p.n_slots;
    const int slot_tiles = p.slot_tiles;
    const int a_exp = p.vec_exp;
    const int n_steps = p.n_steps;
    const int zero = __builtin_amdgcn_readfirstlane(p.zero);   // 0, opaque to the compiler
    // FP32 leaf factors: up to n_dev flushed terms (one without an error model)
    const double leaf_cert = (p.n_dev > 1 ? p.n_dev : 1) * kTau32 * kTwoPowMinusMatExp;
    const double two_m_a = ldexp(1.0, -a_exp);

    // dynamic LDS: scaled-path bookkeeping per slot, the tile's leaf counts,
    // then LDS partials.  Per slot and family (FW of them):
    //   m_cert  W == 1: the stored vector's flush bound; W > 1: before renormalising
    //   m_esum  W == 1: its exponent E (true = stored 2^(E + bsh)); W > 1: e_sum of the raw product
    //   m_bsh   W == 1: its renormalisation shift a - e
    //   m_wmax  W > 1: per-wave maxima;  m_l2s  W > 1: log2 of the nominal scale (per slot)
    extern __shared__ __align__(32) unsigned char s_dyn[];
    double* m_cert = (double*)s_dyn;
    float* m_wmax = (float*)(m_cert + n_slots * FW);
    int* m_esum = (int*)(m_wmax + n_slots * W * FW);
    int* m_bsh = m_esum + n_slots * FW;
    int* m_l2s = m_bsh + n_slots * FW;
    const int meta = mfma_meta_bytes(n_slots, W, SCALED, FB);
    int* s_cnt = (int*)(s_dyn + meta);                // [leaf][FW]
    V4* frag = LDSP ? (V4*)(s_dyn + meta + mfma_count_bytes(p.n_leaves, FB))
                    : (V4*)p.scratch + (long long)blockIdx.x * n_slots * slot_tiles * FB * 64;
    __shared__ double s_rkey[W][FW], s_rx[W][FW], s_rxmax[W][FW];
    __shared__ int s_rj[W][FW];

    // Static, XCD-aware split of the class's tiles: workgroup b runs on XCD
    // b % 8 (dispatch order), each XCD takes one contiguous range of the list
    // (consecutive tiles share a parameter point, hence their matrices, in
    // that XCD's L2) and its workgroups stride through it.  Tiles of one class
    // cost the same, so no work counter is needed.
    const int n_tiles = *as_const(p.n_tiles_dev);
    const int groups = (gridDim.x % 8 == 0) ? 8 : 1;
    const int per_group = gridDim.x / groups;
    const int grp = blockIdx.x % groups;
    const int t_lo = (int)((long long)n_tiles * grp / groups);
    const int t_hi = (int)((long long)n_tiles * (grp + 1) / groups);
    for (int tile = t_lo + (int)blockIdx.x / groups; tile < t_hi; tile += per_group) {
        const TileDesc td = ld_const(p.tiles, tile);
        const BucketDev bk = ld_const(p.buckets, td.bucket);
        const int nf = td.nf < 0 ? -td.nf : td.nf;
        // the tile's leaf counts -> LDS (lane fl of block f holds family 16 f + fl)
        for (int i = threadIdx.x; i < p.n_leaves * FW; i += 64 * W) {
            const int f = i & (FW - 1);
            const int ff = f < nf ? f : 0;
            const int fam = td.nf < 0 ? p.fam_list[td.fam0 + ff] : bk.fam_begin + td.fam0 + ff;
            s_cnt[i] = p.counts_T[(long long)(i / FW) * p.F + fam];
        }
        __syncthreads();   // counts visible
        bool valid[FB];
        int dst[FB];
#pragma unroll
        for (int f = 0; f < FB; ++f) {
            valid[f] = 16 * f + fl < nf;
            const int fi = valid[f] ? 16 * f + fl : 0;
            const int fs = td.nf < 0 ? p.fam_list[td.fam0 + fi] : bk.fam_begin + td.fam0 + fi;
            dst[f] = p.perm[fs];   // the caller's index of this lane's family of block f
        }
        const __attribute__((address_space(4))) int* nslot = as_const(p.node_slot + (long long)td.point * p.n_nodes);
        const int nts = bk.Lp / kTile;        // k-tiles of every child vector
        const int rows_root = bk.Rp / kTile;  // root rows are sizes 0..R
        // flush-bound terms of an interior factor (FP32 matrix entries, MFMA
        // products and stored vector entries that may flush)
        const double cint = SCALED ? kTau32 * (bk.Lp * kTwoPowMinusMatExp * (1.0 + two_m_a) + two_m_a) : 0.0;

        // per-lane offsets of this wave's output tiles, non-root [0] / root [1]:
        // A operand row (FP32 row-major) or column block (FP64 c-major), and
        // leaf rows within a c-major column
        // (separate arrays, chosen with selects: an array indexed by the
        // root flag would be placed in scratch memory)
        int offA0[NTW], offA1[NTW], offL0[NTW], offL1[NTW];
        // FP32: this wave's run of row tiles, non-root [c0, c1) / root [r0, r1)
        const int c0 = kF64 ? 0 : wv * nts / W, c1 = kF64 ? 0 : (wv + 1) * nts / W;
        const int r0 = kF64 ? 0 : wv * rows_root / W, r1 = kF64 ? 0 : (wv + 1) * rows_root / W;
#pragma unroll
        for (int i = 0; i < NTW; ++i) {
            const int u0 = kF64 ? wv + W * i : c0 + i, e0 = kF64 ? nts : c1;
            const int u1 = kF64 ? wv + W * i : r0 + i, e1 = kF64 ? rows_root : r1;
            const int t0 = u0 < e0 ? u0 : nts - 1;
            const int t1 = u1 < e1 ? u1 : rows_root - 1;
            offA0[i] = kF64 ? 4 * g * Np + kTile * t0 + my_arow : (kTile * t0 + my_arow) * Np + 4 * g;
            offA1[i] = kF64 ? 4 * g * Np + kTile * t1 + my_arow : (kTile * t1 + my_arow) * Np + 4 * g;
            offL0[i] = kTile * t0 + 4 * g;
            offL1[i] = kTile * t1 + 4 * g;
        }
        // one-wave FP32 kernels start a non-root product from a 0/1 row mask
        // (rows past the bound are 0 and stay 0), so the node's store needs no
        // masking; the others mask when they store
        constexpr bool kMaskFirst = SCALED && W == 1 && NTW == 2 && FB == 1;   // registers to spare: the 2-tile class
        V4 keep[kMaskFirst ? NTW : 1];
#pragma unroll
        for (int i = 0; i < (kMaskFirst ? NTW : 1); ++i)
#pragma unroll
            for (int r = 0; r < 4; ++r) keep[i][r] = kTile * (wv + W * i) + 4 * g + r < bk.L ? T(1) : T(0);

        // loads of step st (clamped to the last step) into pf; its descriptor
        // into sd.  A leaf step loads its column for family block k into
        // pf[k] (FB = 2), or the same column twice (FB = 1).
        // k-tiles [klo, klo + kn) of an interior step: the union of the bands
        // of the wave's live row tiles (at least one k-tile: a wave whose
        // rows are all 0 multiplies one zero block).  A one-wave workgroup
        // holds every row tile, whose union is the whole vector: no lookup.
        const __attribute__((address_space(4))) int* band = as_const(p.band);
        auto kband = [&](int slot, bool root, int& klo, int& kn) {
            klo = 0;
            kn = nts;
            if constexpr (!kF64 && W > 1) {
                if (!band) return;
                const int a = root ? r0 : c0, b = root ? r1 : c1;
                int lo = nts, hi = 0;
#pragma unroll
                for (int i = 0; i < NTW; ++i)
                    if (a + i < b) {
                        const int e = band[(long long)slot * p.band_stride + a + i];
                        lo = lo < (e & 0xffff) ? lo : (e & 0xffff);
                        hi = hi > (e >> 16) ? hi : (e >> 16);
                    }
                hi = hi < nts ? hi : nts;
                if (lo >= hi) {
                    lo = 0;
                    hi = 1;
                }
                klo = lo;
                kn = hi - lo;
            }
        };
        auto prefetch = [&](V4 (&pf)[PK][NTW], StepDev& sd, int& klo, int& kn, int st) {
            // wave-uniform index: the descriptor comes through the scalar
            // cache, so waiting for it does not wait for vector loads
            st = __builtin_amdgcn_readfirstlane(st < n_steps ? st : n_steps - 1);
            sd = ld_const(p.steps, st);
            const bool root = (sd.ctl & kStepRoot) != 0;
            const bool leaf = sd.leaf_col >= 0;
            const int slot = nslot[sd.node];
            const T* __restrict__ M = (leaf ? lpool : pool) + (long long)slot * p.slot_elems;
            klo = 0;
            kn = nts;
            if (!leaf) kband(slot, root, klo, kn);
#pragma unroll
            for (int k = 0; k < PK; ++k) {
                const int cn = s_cnt[(leaf ? sd.leaf_col : 0) * FW + (FB > 1 ? 16 * k : 0) + fl] * Np;
#pragma unroll
                for (int i = 0; i < NTW; ++i) {
                    const int oa = root ? offA1[i] : offA0[i];
                    const int ol = cn + (root ? offL1[i] : offL0[i]);
                    const int kk = klo + (k < kn ? k : kn - 1);
                    if constexpr (!kF64) {
                        const int off = leaf ? ol + (FB > 1 ? 0 : 4 * k * zero) : oa + kTile * kk;
                        pf[k][i] = *(const floatx4*)(M + off);
                    } else {
                        const int base = leaf ? ol : oa + kTile * kk * Np;
                        const int step = leaf ? 1 + zero : Np;
                        pf[k][i] = V4{M[base], M[base + step], M[base + 2 * step], M[base + 3 * step]};
                    }
                }
            }
        };

        // per-op state (SCALED bookkeeping: a renormalised child v' holds
        // v_c = v' 2^E_c with max v' < 2^a; its factor X' = M' v'
        // (M' = 2^kMatExp M) holds X_c = X' 2^(E_c - kMatExp) and is at most
        // 2^(kMatExp + a), leaf factors 2^kMatExp.
        //   e_sum  sum_c (E_c - kMatExp): true product = raw 2^e_sum (per family)
        //   log2S  log2 of the product's nominal scale prod_c 2^(kMatExp + a_c)
        //   cert   sum_c of the factors' relative flush errors (w.r.t. their
        //          nominal scale): inherited bound + what flushing can drop
        //          (per family)
        V4 acc[FB][NTW];
        int e_sum[FB], log2S = 0;
        double cert[FB];
#pragma unroll
        for (int f = 0; f < FB; ++f) {
            e_sum[f] = 0;
            cert[f] = 0.0;
        }

        auto consume = [&](const V4 (&pf)[PK][NTW], const StepDev& sd, int klo, int kn) {
            const bool root = (sd.ctl & kStepRoot) != 0;
            const int rows = root ? rows_root : nts;
            const int out_slot = sd.ctl & kStepSlotMask;
            int tt[NTW];
            bool live[NTW];
#pragma unroll
            for (int i = 0; i < NTW; ++i) {
                const int u = kF64 ? wv + W * i : (root ? r0 : c0) + i;
                live[i] = u < (kF64 ? rows : root ? r1 : c1);
                tt[i] = live[i] ? u : rows - 1;
            }
            if (sd.ctl & kStepFirst) {
#pragma unroll
                for (int f = 0; f < FB; ++f) {
#pragma unroll
                    for (int i = 0; i < NTW; ++i) {
                        if constexpr (kMaskFirst) acc[f][i] = root ? V4{1, 1, 1, 1} : keep[i];
                        else acc[f][i] = V4{1, 1, 1, 1};
                    }
                    e_sum[f] = 0;
                    cert[f] = 0.0;
                }
                log2S = 0;
            }
            V4 x[FB][NTW];
            if (sd.leaf_col >= 0) {
                if (SCALED) {
#pragma unroll
                    for (int f = 0; f < FB; ++f) {
                        e_sum[f] -= kMatExp;
                        cert[f] += leaf_cert;
                    }
                    log2S += kMatExp;
                }
#pragma unroll
                for (int f = 0; f < FB; ++f)
#pragma unroll
                    for (int i = 0; i < NTW; ++i) x[f][i] = pf[f][i];
                // The leaf step ignores the other prefetched k-tiles; taking
                // them here makes both paths leave no load of this buffer in
                // flight, so its next prefetch needs no wait on the other one.
#pragma unroll
                for (int k = FB; k < PK; ++k)
#pragma unroll
                    for (int i = 0; i < NTW; ++i) await_loaded(pf[k][i]);
            } else {
                int bsh[FB];
#pragma unroll
                for (int f = 0; f < FB; ++f) bsh[f] = 0;
                if (SCALED) {
                    log2S += kMatExp + a_exp;
#pragma unroll
                    for (int f = 0; f < FB; ++f) {
                        const int q = sd.in_slot * FW + 16 * f + fl;
                        if (W == 1) {
                            bsh[f] = m_bsh[q];
                            e_sum[f] += m_esum[q] - kMatExp;
                            cert[f] += m_cert[q] + cint;
                        } else {
                            float m = 0.f;
                            for (int w2 = 0; w2 < W; ++w2)
                                m = fmaxf(m, m_wmax[(sd.in_slot * W + w2) * FW + 16 * f + fl]);
                            int e = 0;
                            const float mant = m > 0.f ? frexpf(m, &e) : 1.f;
                            bsh[f] = a_exp - e;
                            e_sum[f] += m_esum[q] + e - a_exp - kMatExp;
                            cert[f] += (m > 0.f ? (ldexp(m_cert[q] * 1.01, m_l2s[sd.in_slot] - e) + ldexp(kTau32, -e)) *
                                                      rcp_up(mant)
                                                : 1e300) +
                                       cint;
                        }
                    }
                }
                // FP32: the step's k-tiles start at klo (column 16 klo of the
                // row-major rows, fragment klo of the child)
                const T* __restrict__ M = pool + (long long)nslot[sd.node] * p.slot_elems + (kF64 ? 0 : kTile * klo);
                int ao[NTW];
#pragma unroll
                for (int i = 0; i < NTW; ++i) {
#pragma unroll
                    for (int f = 0; f < FB; ++f) x[f][i] = V4{0, 0, 0, 0};
                    ao[i] = root ? offA1[i] : offA0[i];
                }
                const V4* bsrc = frag + ((long long)sd.in_slot * slot_tiles + klo) * FB * 64 + lane;
                child_gemm<T, NTW, Ring<T, NTW, W, FB>::D, SCALED, FB>(x, M, Np, ao, pf, bsrc, kn, bsh);
            }
#pragma unroll
            for (int f = 0; f < FB; ++f)
#pragma unroll
                for (int i = 0; i < NTW; ++i) acc[f][i] *= x[f][i];
            if (!(sd.ctl & kStepLast)) return;

            if (!root) {
                T wmax[FB];
#pragma unroll
                for (int f = 0; f < FB; ++f) {
                    wmax[f] = 0;
#pragma unroll
                    for (int i = 0; i < NTW; ++i) {
                        if (!live[i]) continue;
                        V4 v = acc[f][i];
                        if constexpr (!kMaskFirst)
#pragma unroll
                            for (int r = 0; r < 4; ++r)
                                if (kTile * tt[i] + 4 * g + r >= bk.L) v[r] = 0;   // rows past the bound
                        frag[(((long long)out_slot * slot_tiles + tt[i]) * FB + f) * 64 + lane] = v;
                        if (SCALED)
#pragma unroll
                            for (int r = 0; r < 4; ++r) {
                                if constexpr (W == 1) wmax[f] = fmax(wmax[f], v[r]);   // entries are >= 0
                                else wmax[f] = wmax[f] > v[r] ? wmax[f] : v[r];
                            }
                    }
                }
                if (SCALED) {
#pragma unroll
                    for (int f = 0; f < FB; ++f) {
                        // maximum over the family's four lanes fl, fl+16, fl+32,
                        // fl+48 (the one-wave kernels use lane swaps; the 4-tile
                        // multi-wave shapes keep shuffles, which leave them the
                        // registers for two waves per SIMD)
                        float m = (float)wmax[f];
                        if constexpr (W == 1) {
                            m = fmaxf(m, __builtin_bit_cast(float, __builtin_amdgcn_ds_swizzle(
                                                                       __builtin_bit_cast(int, m), kSwizzleXor16)));
                            m = xor32_max(m);
                        } else {
                            m = fmaxf(m, __shfl_xor(m, 16));
                            m = fmaxf(m, __shfl_xor(m, 32));
                        }
                        const int q = out_slot * FW + 16 * f + fl;
                        if (W == 1) {
                            // finish the stored vector: shift, exponent, flush
                            // bound re-expressed relative to the new scale (it
                            // grows by S / m, the factor by which the product
                            // fell short of its nominal scale)
                            int e = 0;
                            const float mant = m > 0.f ? frexpf(m, &e) : 1.f;
                            if (g == 0) {
                                m_bsh[q] = a_exp - e;
                                m_esum[q] = e_sum[f] + e - a_exp;
                                m_cert[q] = m > 0.f ? (ldexp(cert[f] * 1.01, log2S - e) + ldexp(kTau32, -e)) *
                                                          rcp_up(mant)
                                                    : 1e300;
                            }
                        } else {
                            if (g == 0) m_wmax[(out_slot * W + wv) * FW + 16 * f + fl] = m;
                            if (wv == 0 && g == 0) {
                                m_esum[q] = e_sum[f];
                                m_cert[q] = cert[f];
                            }
                        }
                    }
                    if (W > 1 && wv == 0 && lane == 0) m_l2s[out_slot] = log2S;
                }
                // the node's vector is complete before anyone reads it (one
                // wave: LDS keeps its own order, only the compiler must not
                // move accesses across)
                if (!LDSP) __syncthreads();
                else if (W > 1) lds_barrier();
                else asm volatile("" ::: "memory");
                return;
            }

            // root: best x_j * prior_j over this lane's rows (row = root size
            // j + 1, j = 0..R-1), then across lanes and waves
            const double* __restrict__ pr = p.prior64 + bk.prior_off;
            double pv[NTW][4];
#pragma unroll
            for (int i = 0; i < NTW; ++i)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int j = kTile * tt[i] + 4 * g + r - 1;
                    pv[i][r] = pr[j < 0 ? 0 : (j < bk.R ? j : bk.R - 1)];
                }
            // every prior arrives here, whichever rows use it, so no load of
            // them is left in flight when the next tile's prefetch reuses
            // their registers (that would stall every step's prefetch)
#pragma unroll
            for (int i = 0; i < NTW; ++i)
#pragma unroll
                for (int r = 0; r < 4; ++r) asm volatile("" ::"v"(pv[i][r]));
            RootBest rb[FB];
#pragma unroll
            for (int f = 0; f < FB; ++f) {
                rb[f] = RootBest{-__builtin_huge_val(), 0.0, -1, 0.0};
#pragma unroll
                for (int i = 0; i < NTW; ++i) {
                    if (!live[i]) continue;
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const int j = kTile * tt[i] + 4 * g + r - 1;
                        if (j < 0 || j >= bk.R) continue;
                        const double xv = (double)acc[f][i][r];
                        if (p.root_out && valid[f])
                            p.root_out[(long long)dst[f] * p.root_stride + j] = SCALED ? ldexp(xv, e_sum[f]) : xv;
                        root_take(rb[f], xv * pv[i][r], xv, j);
                    }
                }
                rb[f] = root_combine(rb[f]);
            }
            if (W > 1) {
#pragma unroll
                for (int f = 0; f < FB; ++f)
                    if (g == 0) {
                        s_rkey[wv][16 * f + fl] = rb[f].key;
                        s_rx[wv][16 * f + fl] = rb[f].x;
                        s_rxmax[wv][16 * f + fl] = rb[f].xmax;
                        s_rj[wv][16 * f + fl] = rb[f].j;
                    }
                lds_barrier();
                if (wv == 0)
#pragma unroll
                    for (int f = 0; f < FB; ++f)
                        for (int w2 = 1; w2 < W; ++w2)
                            root_merge(rb[f], s_rkey[w2][16 * f + fl], s_rx[w2][16 * f + fl], s_rj[w2][16 * f + fl],
                                       s_rxmax[w2][16 * f + fl]);
            }
            if (wv == 0) {
#pragma unroll
                for (int f = 0; f < FB; ++f) {
                    double val;
                    int flag = 0;
                    if (SCALED) {
                        // rb.key = x_j * prior_j in units of 2^e_sum
                        if (!(rb[f].key > 0)) {
                            flag = kFlagRecompute64;
                            val = 0.0;
                        } else {
                            const double log_scale = (double)e_sum[f] * kLn2;
                            if (p.model == 0)
                                val = log(rb[f].x) + log(p.prior64[bk.prior_off + rb[f].j]) + log_scale;
                            else
                                val = ldexp(rb[f].key, e_sum[f]);
                            // flush bound (absolute, in units of 2^e_sum) against
                            // the likelihood, and FP64 range of the unscaled values
                            const double bound = (ldexp(cert[f], log2S) + kTau32) * 1.01;
                            if (!(bound <= kCertEta * rb[f].key)) flag |= kFlagRecompute64;
                            if (!(log(rb[f].xmax) + log_scale >= log(kFp64Floor))) flag |= kFlagRecompute64;
                            if (!(log(rb[f].key) + log_scale >= log(kFp64Floor))) flag |= kFlagRecompute64;
                        }
                    } else {
                        // base: log(x_j prior_j) at the best j (base_model.cpp:86
                        // takes log x_j + log prior_j); gamma: x_j prior_j
                        val = p.model == 0 ? log(rb[f].key) : rb[f].key;
                        if (rb[f].xmax == 0.0) flag |= kFlagZeroSum;
                        if (rb[f].xmax < kFp64Floor || rb[f].key < kFp64Floor) flag |= kFlagRecomputePlain;
                    }
                    if (valid[f] && g == 0) {
                        p.out_val[(long long)td.point * p.F + dst[f]] = val;
                        p.out_flag[(long long)td.point * p.F + dst[f]] = flag;
                    }
                }
            }
        };

        V4 pfa[PK][NTW], pfb[PK][NTW];
        StepDev sda, sdb;
        int kla, kna, klb, knb;
        prefetch(pfa, sda, kla, kna, 0);
        for (int st = 0; st < n_steps; st += 2) {
            prefetch(pfb, sdb, klb, knb, st + 1);
            consume(pfa, sda, kla, kna);
            if (st + 1 < n_steps) {
                prefetch(pfa, sda, kla, kna, st + 2);
                consume(pfb, sdb, klb, knb);
            }
        }
        // the last prefetch (past the end) is never consumed: take it here, so
        // the next tile's step loop starts with no load in flight
#pragma unroll
        for (int k = 0; k < PK; ++k)
#pragma unroll
            for (int i = 0; i < NTW; ++i) {
                await_loaded(pfa[k][i]);
                await_loaded(pfb[k][i]);
            }
        __syncthreads();   // every wave is done with this tile's LDS before the next one
    }
}

// ---------------------------------------------------------------------------
// Launcher: one kernel per (tiles per wave, waves per workgroup, partials in
// LDS) shape; the host picks the shape per size class.
// ---------------------------------------------------------------------------
// The kernel shapes (tiles per wave, waves per workgroup, partials in LDS);
// (3|4, 16) cover vectors of up to 64 tiles (1024 rows) in one pass.
#define CAFE_MFMA_SHAPES(X, T, S) \
    X(T, 1, 1, S, true) X(T, 1, 1, S, false) X(T, 2, 1, S, true) X(T, 2, 1, S, false) \
    X(T, 3, 1, S, true) X(T, 3, 1, S, false) X(T, 4, 1, S, true) X(T, 4, 1, S, false) \
    X(T, 3, 2, S, true) X(T, 3, 2, S, false) X(T, 4, 2, S, true) X(T, 4, 2, S, false) \
    X(T, 3, 4, S, true) X(T, 3, 4, S, false) X(T, 4, 4, S, true) X(T, 4, 4, S, false) \
    X(T, 3, 8, S, false) X(T, 4, 8, S, false)
#define CAFE_MFMA_WIDE(X, T, S) X(T, 3, 16, S, false) X(T, 4, 16, S, false)
// Latency shapes: one or two output tiles per wave, so that one family
// tile's chain of steps is as short as the workgroup allows (the FP64
// recompute's few tiles; FP32 classes with fewer tiles than resident slots)
#define CAFE_MFMA_CHAIN(X, T, S) \
    X(T, 1, 4, S, true) X(T, 1, 4, S, false) X(T, 1, 8, S, true) X(T, 1, 8, S, false) X(T, 2, 8, S, true) \
    X(T, 2, 8, S, false) X(T, 2, 16, S, false) X(T, 1, 16, S, true) X(T, 1, 16, S, false)

// 32-family tiles (FP32 only: a leaf step's two columns use both prefetch
// buffers) for vectors of up to 32 tiles, whose k-loops stream their
// matrices.  Two output tiles per wave and block hold ~190 VGPRs, four would
// spill; sixteen waves of two (the 17-32-tile vectors) are capped at 128
// VGPRs and keep ~30 values in scratch.
#define CAFE_MFMA_FB2(X, T, S) \
    X(T, 2, 4, S, true) X(T, 2, 4, S, false) X(T, 1, 8, S, true) X(T, 1, 8, S, false) X(T, 2, 8, S, true) \
    X(T, 2, 8, S, false) X(T, 1, 16, S, false) X(T, 2, 16, S, false)

// One instance per shape; with per_cu set, reports how many workgroups of it
// fit on one CU instead of launching.
template <typename T, bool SCALED>
static hipError_t shape_dispatch(const MfmaShape& s, int grid, const PruneParams& p, hipStream_t st, int* per_cu)
{
#define CAFE_MFMA_CASE_FB(T_, NTW, W, S_, LDSP, FB)                                                          \
    if (s.ntw == NTW && s.w == W && s.lds == LDSP && s.fb == FB) {                                          \
        if (per_cu)                                                                                         \
            return hipOccupancyMaxActiveBlocksPerMultiprocessor(                                            \
                per_cu, k_prune_mfma<T, NTW, W, SCALED, LDSP, FB>, 64 * W, s.shm);                          \
        hipLaunchKernelGGL((k_prune_mfma<T, NTW, W, SCALED, LDSP, FB>), dim3(grid), dim3(64 * W), s.shm, st, p); \
        return hipGetLastError();                                                                           \
    }
#define CAFE_MFMA_CASE(T_, NTW, W, S_, LDSP) CAFE_MFMA_CASE_FB(T_, NTW, W, S_, LDSP, 1)
#define CAFE_MFMA_CASE2(T_, NTW, W, S_, LDSP) CAFE_MFMA_CASE_FB(T_, NTW, W, S_, LDSP, 2)
    CAFE_MFMA_SHAPES(CAFE_MFMA_CASE, T, SCALED)
    CAFE_MFMA_WIDE(CAFE_MFMA_CASE, T, SCALED)
    CAFE_MFMA_CHAIN(CAFE_MFMA_CASE, T, SCALED)
    if constexpr (sizeof(T) == 4) {
        CAFE_MFMA_FB2(CAFE_MFMA_CASE2, T, SCALED)
    }
#undef CAFE_MFMA_CASE2
#undef CAFE_MFMA_CASE
#undef CAFE_MFMA_CASE_FB
    return hipErrorInvalidValue;
}

// Workgroups of a shape that fit on one CU at once (registers and LDS), so a
// launch can size its grid to what is resident: the static tile split gives
// every workgroup an equal share, and a workgroup that has to wait for a
// slot would finish its share after all the others.
hipError_t mfma_occupancy(bool fp64, const MfmaShape& s, int* per_cu)
{
    *per_cu = 0;
    const PruneParams none{};
    return fp64 ? shape_dispatch<double, false>(s, 0, none, nullptr, per_cu)
                : shape_dispatch<float, true>(s, 0, none, nullptr, per_cu);
}

hipError_t launch_prune_mfma(bool fp64, const MfmaShape& s, int grid, const PruneParams& p, hipStream_t st)
{
    if (grid < 1) grid = 1;
    return fp64 ? shape_dispatch<double, false>(s, grid, p, st, nullptr)
                : shape_dispatch<float, true>(s, grid, p, st, nullptr);
}

}  // namespace cafe
