// prune_lane.hip -- batched FP32 pruning of small vectors, one family per lane
// (replaces inference_prune / compute_node_probability, src/core.cpp:134-145,
// src/probability.cpp:175-330, and the root reductions of base_model.cpp:77-95
// and gamma_core.cpp:143-165, for the families of the smallest size buckets).
//
// A wave carries 64 families of one size bucket at one parameter point through
// the whole tree.  Lane l owns family l: its partial vectors (sizes 0..LR-1),
// the running product of the current node and the node's factor all live in
// the lane's registers, so a node costs no data movement between lanes and the
// scaled-FP32 bookkeeping (exponent, shift, flush bound) is lane-local.  An
// interior child's factor x = M' v is LR x LR fused multiply-adds per lane
// whose matrix operand is a wave-uniform scalar register: the wave streams
// column k of M' (c-major FP32 pool, entries times 2^kMatExp) through scalar
// loads while v[k] is the lane's own value.  A leaf child's factor is column
// count of M' gathered per lane.  No row is padded beyond LR, against the
// 16-row tiles of the matrix-core kernel (21 rows: 1.0x the work instead of
// 2.3x), and a SIMD holds 64 families per wave.
//
// Partial vectors are in registers, NS arrays of LR floats.  The program
// (post order) puts each node's op right after the op of its last interior
// child, so that child's vector is read straight from the last array, the
// "hot" one, which the op just before wrote (its output is marked kHotSlot,
// as is the child read): no selects.  Only the other interior children (the
// first of two interior siblings) are stored, in arrays 0 .. NS-2 indexed by
// the program's slot numbers through selects under wave-uniform masks and
// weights (per-slot copies of the code under wave-uniform branches run ~6%
// faster alone but hold ~10% more registers, which costs more beside the
// other classes; so does a hot array on top of NS stored ones).  A stored
// output may take the slot of one of its children (the children are all read
// before the output is written, by the same lane).
//
// Numerics follow the single-wave matrix-core kernel exactly: stored vectors
// are renormalised to a maximum just below 2^vec_exp, the exponent is carried
// in an integer, and a bound on what FP32 flushing to zero could have dropped
// decides at the root whether the family is kept or recomputed in FP64.
#include "prune_common.h"

namespace cafe {

// Slot access with a wave-uniform slot number s.  The register arrays are
// indexed by constants only: a vector entry is selected bitwise under uniform
// masks, the scalar bookkeeping is blended with uniform 0/1 weights (stored
// values are finite, so the blend is exact).  Plain selects would be turned
// by the compiler into one load at a variable index, which moves the arrays
// to scratch memory.
template <int NS>
struct SlotWeights {
    float w[NS];
    unsigned m[NS];   // all ones for slot s, else zero
    __device__ __forceinline__ explicit SlotWeights(int s)
    {
#pragma unroll
        for (int c = 0; c < NS; ++c) {
            w[c] = s == c ? 1.f : 0.f;
            m[c] = s == c ? ~0u : 0u;
        }
    }
};
// (m & a) | (~m & b) with a wave-uniform mask (all ones or zero, the same in
// every lane): one bitfield insert per slot past the first.  The asm operand
// keeps the compiler from rewriting the chain as a select at a variable index.
__device__ __forceinline__ float bit_select(unsigned m, float a, float b)
{
    float r;
    asm("v_bfi_b32 %0, %1, %2, %3" : "=v"(r) : "v"(m), "v"(a), "v"(b));
    return r;
}
// (stored slots are arrays 0 .. NSS-1 of the NS > NSS arrays)
template <int NSS, int NS, int LR>
__device__ __forceinline__ float pick(const float (&sv)[NS][LR], const SlotWeights<NSS>& sw, int k)
{
    float r = sv[0][k];
#pragma unroll
    for (int c = 1; c < NSS; ++c) r = bit_select(sw.m[c], sv[c][k], r);
    return r;
}
template <int NSS, int NS, typename V>
__device__ __forceinline__ V pick(const V (&a)[NS], const SlotWeights<NSS>& sw)
{
    V r = 0;
#pragma unroll
    for (int c = 0; c < NSS; ++c) r += sw.w[c] != 0.f ? a[c] : V(0);
    return r;
}
// (the store is a select per element as well: a branch per slot would let
// the compiler merge the stores into one at a variable index, which moves the
// arrays to scratch memory)
template <int NSS, int NS, int LR>
__device__ __forceinline__ void put(float (&sv)[NS][LR], int (&se)[NS], double (&sc)[NS], int out,
                                    const float (&v)[LR], int e, double c)
{
#pragma unroll
    for (int q = 0; q < NSS; ++q) {
        const bool here = out == q;
#pragma unroll
        for (int i = 0; i < LR; ++i) sv[q][i] = here ? v[i] : sv[q][i];
        se[q] = here ? e : se[q];
        sc[q] = here ? c : sc[q];
    }
}

// Three waves per SIMD (168 VGPRs) hold 21-row vectors in up to 3 arrays (two
// stored slots and the hot array); larger shapes ask for two.
template <int LR, int NS> struct LaneWaves { static constexpr int v = LR * (NS + 3) <= 130 ? 3 : 2; };

template <int LR, int NS>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(LaneWaves<LR, NS>::v))) void k_prune_lane(PruneParams p)
{
    const int lane = threadIdx.x;
    const int Np = p.Np;
    const int a_exp = p.vec_exp;
    const int n_steps = p.n_steps;
    const long long slot_elems = p.slot_elems;
    // FP32 leaf factors: up to n_dev flushed terms (one without an error model)
    const double leaf_cert = (p.n_dev > 1 ? p.n_dev : 1) * kTau32 * kTwoPowMinusMatExp;
    const double two_m_a = ldexp(1.0, -a_exp);
    // flush-bound terms of an interior factor relative to its nominal scale:
    // LR matrix entries and products that may flush, and the stored vector's
    // own entries
    const double cint = kTau32 * (LR * kTwoPowMinusMatExp * (1.0 + two_m_a) + two_m_a);
    extern __shared__ int s_cnt[];   // [leaf][64]: the tile's leaf counts, lane l's in column l

    const int n_tiles = *as_const(p.n_tiles_dev);
    const int groups = (gridDim.x % 8 == 0) ? 8 : 1;   // XCD-aware static split (see k_prune_mfma)
    const int per_group = gridDim.x / groups;
    const int grp = blockIdx.x % groups;
    const int t_lo = (int)((long long)n_tiles * grp / groups);
    const int t_hi = (int)((long long)n_tiles * (grp + 1) / groups);
    for (int tile = t_lo + (int)blockIdx.x / groups; tile < t_hi; tile += per_group) {
        const TileDesc td = ld_const(p.tiles, tile);
        const BucketDev bk = ld_const(p.buckets, td.bucket);
        const int nf = td.nf < 0 ? -td.nf : td.nf;
        const bool valid = lane < nf;
        const int fi = valid ? lane : 0;
        const int fs = td.nf < 0 ? p.fam_list[td.fam0 + fi] : bk.fam_begin + td.fam0 + fi;
        const int dst = p.perm[fs];   // the caller's index of this lane's family
        for (int l = 0; l < p.n_leaves; ++l) s_cnt[l * 64 + lane] = p.counts_T[(long long)l * p.F + fs];
        const __attribute__((address_space(4))) int* nslot = as_const(p.node_slot + (long long)td.point * p.n_nodes);

        constexpr int H = NS - 1;   // the hot array; stored slots 0 .. NS-2
        float sv[NS][LR];   // partial vectors (renormalised)
        int se[NS];         // their exponents: true vector = stored 2^se
        double sc[NS];      // their flush bounds, relative to their maximum
        float acc[LR];      // the current node's running product
        int e_sum = 0, log2S = 0;
        double cert = 0.0;
#pragma unroll
        for (int c = 0; c < NS; ++c) {
            se[c] = 0;
            sc[c] = 0.0;
#pragma unroll
            for (int i = 0; i < LR; ++i) sv[c][i] = 0.f;
        }
#pragma unroll
        for (int i = 0; i < LR; ++i) acc[i] = 1.f;

        // node by node (the program's ops, root last): the children's factors
        // into acc, then the node is stored in its slot.  Every non-root
        // node ends in the same store, with no branch around it: a branch
        // makes the compiler keep copies of all slots.
        int st = 0;
        for (;;) {
            bool root = false;
            int out = 0;
            for (;;) {
                const StepDev sd = ld_const(p.steps, st++);
                root = (sd.ctl & kStepRoot) != 0;
                out = sd.ctl & kStepSlotMask;
                // the node's first child sets the product (1 x factor, exactly),
                // later children multiply into it
                const bool first = (sd.ctl & kStepFirst) != 0;
                if (first) {
                    e_sum = 0;
                    log2S = 0;
                    cert = 0.0;
                }
                const long long base = (long long)nslot[sd.node] * slot_elems;
                if (sd.leaf_col >= 0) {
                    // leaf: column `count` of the (error-model folded) matrix
                    const int cnt = s_cnt[sd.leaf_col * 64 + lane];
                    const float* __restrict__ col = p.leaf32 + base + (long long)cnt * Np;
                    e_sum -= kMatExp;
                    log2S += kMatExp;
                    cert += leaf_cert;
                        // 16-byte loads: a column starts 128-byte aligned (Np is a
                    // multiple of 32) and holds Np >= 4 ceil(LR / 4) entries
                    // (Np >= 32 always; 33-row classes have N >= 22, Np >= 64)
                    typedef float f4 __attribute__((ext_vector_type(4)));
                    auto gather = [&](float (&lv)[LR]) {
#pragma unroll
                        for (int q = 0; q < (LR + 3) / 4; ++q) {
                            const f4 c4 = *(const f4*)(col + 4 * q);
#pragma unroll
                            for (int r = 0; r < 4; ++r)
                                if (4 * q + r < LR) lv[4 * q + r] = c4[r];
                        }
                    };
                    // a first child's factor is gathered into the product itself
                    if (first) {
                        gather(acc);
                    } else {
                        float lv[LR];
                        gather(lv);
#pragma unroll
                        for (int i = 0; i < LR; ++i) acc[i] *= lv[i];
                    }
                } else {
                    // interior child, hot or in slot s: x = M' v, column k of M'
                    // from the scalar cache, v[k] the lane's own
                    const __attribute__((address_space(4))) float* M = as_const(p.cm32 + base);
                    float x[LR];
                    // columns through scalar-register buffers: the loads of the
                    // next D columns are in flight while column k is used
                    // (D = 2 for 21 rows, 1 for larger vectors: SGPR budget)
                    constexpr int D = LR <= 24 ? 2 : 1;
                    auto contract = [&](float (&x)[LR], auto vread) {
                        float cbuf[D + 1][LR];
#pragma unroll
                        for (int d = 0; d < D; ++d)
#pragma unroll
                            for (int i = 0; i < LR; ++i) cbuf[d][i] = d < LR ? M[d * Np + i] : 0.f;
#pragma unroll
                        for (int k = 0; k < LR; ++k) {
                            float (&cur)[LR] = cbuf[k % (D + 1)];
                            float (&nxt)[LR] = cbuf[(k + D) % (D + 1)];
                            if (k + D < LR) {
#pragma unroll
                                for (int i = 0; i < LR; ++i) nxt[i] = M[(k + D) * Np + i];
                            }
                            const float vk = vread(k);
                            // column 0 starts the sums with a product: entries
                            // are finite and >= 0, so this equals fma(., ., +0)
#pragma unroll
                            for (int i = 0; i < LR; ++i)
                                x[i] = k == 0 ? cur[i] * vk : __builtin_fmaf(cur[i], vk, x[i]);
                            __builtin_amdgcn_sched_barrier(0);   // no loads further ahead than column k + D
                        }
                    };
                    // (summing a first child straight into the product takes four
                    // copies of the contraction, whose registers spill)
                    int ce;
                    double cc;
                    if (sd.in_slot == kHotSlot) {
                        contract(x, [&](int k) { return sv[H][k]; });
                        ce = se[H];
                        cc = sc[H];
                    } else {
                        const SlotWeights<NS - 1> sw(sd.in_slot);
                        contract(x, [&](int k) { return pick<NS - 1>(sv, sw, k); });
                        ce = pick<NS - 1>(se, sw);
                        cc = pick<NS - 1>(sc, sw);
                    }
                    e_sum += ce - kMatExp;
                    cert += cc + cint;
                    log2S += kMatExp + a_exp;
                    if (first) {
#pragma unroll
                        for (int i = 0; i < LR; ++i) acc[i] = x[i];
                    } else {
#pragma unroll
                        for (int i = 0; i < LR; ++i) acc[i] *= x[i];
                    }
                }
                // a non-root product keeps rows past the bound at 0 (they
                // stay 0); root rows are sizes 0..R
                if (first && !root && bk.L < LR) {
#pragma unroll
                    for (int i = 0; i < LR; ++i) acc[i] = i < bk.L ? acc[i] : 0.f;
                }
                if ((sd.ctl & kStepLast) || st >= n_steps) break;
            }
            if (root || st >= n_steps) break;
            // store the node: shift to a maximum just below 2^a, exponent,
            // flush bound re-expressed relative to the new maximum (it grows
            // by S / m, the factor by which the product fell short of its
            // nominal scale S = 2^log2S)
            float m = 0.f;
#pragma unroll
            for (int i = 0; i < LR; ++i) m = fmaxf(m, acc[i]);   // entries are >= 0
            int e = 0;
            const float mant = m > 0.f ? frexpf(m, &e) : 1.f;
            const int sh = a_exp - e;
            const double cnew =
                m > 0.f ? (ldexp(cert * 1.01, log2S - e) + ldexp(kTau32, -e)) * rcp_up(mant) : 1e300;
            // the output goes to the hot array (the op it would have fed there
            // has read it), and from there to its slot when a later op reads it
#pragma unroll
            for (int i = 0; i < LR; ++i) sv[H][i] = ldexpf(acc[i], sh);
            se[H] = e_sum + e - a_exp;
            sc[H] = cnew;
            if (out != kHotSlot) put<NS - 1>(sv, se, sc, out, sv[H], se[H], sc[H]);
        }
        {
            // root: best x_j * prior_j over root sizes j + 1, j = 0..R-1
            const __attribute__((address_space(4))) double* pr = as_const(p.prior64 + bk.prior_off);
            RootBest rb{-__builtin_huge_val(), 0.0, -1, 0.0};
#pragma unroll
            for (int j = 0; j + 1 < LR; ++j) {
                if (j >= bk.R) continue;
                const double xv = (double)acc[j + 1];
                if (p.root_out && valid) p.root_out[(long long)dst * p.root_stride + j] = ldexp(xv, e_sum);
                root_take(rb, xv * pr[j], xv, j);
            }
            double val;
            int flag = 0;
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
                // flush bound (absolute, in units of 2^e_sum) against the
                // likelihood, and FP64 range of the unscaled values
                const double bound = (ldexp(cert, log2S) + kTau32) * 1.01;
                if (!(bound <= kCertEta * rb.key)) flag |= kFlagRecompute64;
                if (!(log(rb.xmax) + log_scale >= log(kFp64Floor))) flag |= kFlagRecompute64;
                if (!(log(rb.key) + log_scale >= log(kFp64Floor))) flag |= kFlagRecompute64;
            }
            if (valid) {
                p.out_val[(long long)td.point * p.F + dst] = val;
                p.out_flag[(long long)td.point * p.F + dst] = flag;
            }
        }
    }
}

// Rows per lane the kernel instances cover (vectors of sizes 0..B and root
// sizes 0..R with max(B, R) + 1 <= rows): the buckets of bound 20 and 32.
int lane_rows_for(int rows)
{
    if (rows <= 21) return 21;
    if (rows <= 33) return 33;
    return 0;
}

// Largest array count the instances of a row count cover (stored slots + the
// hot array).
int lane_max_slots(int lane_rows) { return lane_rows == 21 ? 5 : lane_rows == 33 ? 4 : 0; }

// One instance per (rows, slots); with per_cu set, reports how many
// workgroups of it fit on one CU instead of launching.
hipError_t lane_dispatch(int lr, int ns, int grid, size_t shm, const PruneParams& p, hipStream_t st, int* per_cu)
{
#define CAFE_LANE_CASE(LR, NS)                                                                             \
    if (lr == LR && ns <= NS) {                                                                            \
        if (per_cu) return hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, k_prune_lane<LR, NS>, 64, shm); \
        hipLaunchKernelGGL((k_prune_lane<LR, NS>), dim3(grid), dim3(64), shm, st, p);                      \
        return hipGetLastError();                                                                          \
    }
    CAFE_LANE_CASE(21, 2)
    CAFE_LANE_CASE(21, 3)
    CAFE_LANE_CASE(21, 4)
    CAFE_LANE_CASE(21, 5)
    CAFE_LANE_CASE(33, 2)
    CAFE_LANE_CASE(33, 3)
    CAFE_LANE_CASE(33, 4)
#undef CAFE_LANE_CASE
    return hipErrorInvalidValue;
}

}  // namespace cafe
