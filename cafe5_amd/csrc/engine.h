// engine.h -- device data structures shared by the HIP kernels and the host
// planner of libcafe_gpu.so.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace cafe {

// Matrix pools: element (s, c) = P(s -> c) of slot k is, in the FP64 pool,
// at mt64[k * slot_elems + c * Np + s] ("c-major", so a wave reading rows
// s = lane.. of one column is coalesced); in the FP32 pool at
// mt32[k * slot_elems + s * Np + c] (row-major, so one 16-byte load gives an
// MFMA lane four consecutive contraction entries); and in the FP32 leaf pool
// at lf32[k * slot_elems + c * Np + s] (c-major: one 16-byte load gives a lane
// four consecutive rows of a leaf column).  Np = round_up(N + 16, 32) leaves
// room for 16-row tiles (Np >= round_up(N, 16) + 16); padding is 0.
constexpr int kTile = 16;
constexpr int kMaxSlots = 24;         // live partial vectors per family tile
constexpr int kLgammaN = 4096;        // log-gamma table (integers < 4096)
// The FP32 copy of every matrix is stored multiplied by 2^kMatExp, and FP32
// partial vectors are renormalised to a max just below 2^vec_exp, so that the
// whole FP32 exponent range is used: products of k children stay below 2^127
// when k (kMatExp + vec_exp) <= 126 (the host picks vec_exp from the tree's
// largest child count) while entries far below each vector's max survive.
constexpr int kMatExp = 27;
// Band entry of a row tile: k-tiles [lo, hi) hold its nonzero entries.
__host__ __device__ inline int band_entry(int lo, int hi) { return lo | hi << 16; }
// FP32 copies a key's users read (a plan's build writes only these):
// kCopyRow the row-major pool (interior branches on the matrix cores and the
// split-FP16 halves), kCopyCol the c-major pool (leaf columns), kCopyColTop
// its entries of sizes below the per-lane kernel's rows (interior branches
// of the per-lane classes).
enum : unsigned char { kCopyRow = 1, kCopyCol = 2, kCopyColTop = 4 };
// The split-FP16 pools (hi and lo halves, row-major, see prune_f16.hip) hold
// every matrix times 2^kMatExp16: entries <= 1 stay below FP16's largest
// finite value.
constexpr int kMatExp16 = 15;

// One unit of work of the pruning kernels: families fam0 .. fam0+nf-1 of a
// size bucket (bucket-sorted order) at one parameter point.  nf < 0 marks a
// packed tile: its -nf families are fam_list[fam0 .. fam0-nf-1].
struct TileDesc {
    int bucket;
    int point;
    int fam0;
    int nf;
};

// (bucket, point) run of tiles, the unit in which flagged families are packed
// into recompute tiles; cls is the size class whose lists receive them.
// A chunk of one (bucket, point) segment for the recompute packing:
// families [first, first + kSegChunk) of the bucket at that point.
constexpr int kSegChunk = 1024;
struct SegDesc {
    int bucket;
    int point;
    int cls;
    int first;
};

struct BucketDev {
    int B, R;          // max_family_size, max_root_family_size
    int L, Lp;         // B+1 and its 16-multiple
    int Rp;            // root rows 0..R (size 0 unused) rounded up to 16
    int fam_begin;     // first family (bucket-sorted order)
    int count;
    int prior_off;     // offset of this bucket's prior in the prior tables
};

// Internal node of the evaluation program (Sethi-Ullman order, root last).
struct OpDev {
    int node;
    int out_slot;
    int n_child;
    int child_begin;
};

struct ChildDev {
    int node;       // child node (its branch matrix is node_slot[point][node])
    int leaf_col;   // >= 0: leaf column in the count table; -1: interior
    int in_slot;    // interior: slot holding its partial vector
    int pad;
};

// One child contraction of the evaluation program, flattened in op order
// (the MFMA kernel walks these and prefetches the next one's operands).
struct StepDev {
    int node;       // child node (branch matrix node_slot[point][node])
    int leaf_col;   // >= 0: leaf column; -1: interior child
    int in_slot;    // interior: slot holding its partial vector
    int ctl;        // out_slot | kStepFirst | kStepLast | kStepRoot
};
enum : int {
    kStepFirst = 1 << 8,   // first child of its op: the product starts at 1
    kStepLast = 1 << 9,    // last child: the op's vector is complete
    kStepRoot = 1 << 10,   // the op is the root
    kStepSlotMask = 0xff,
    // per-lane kernel: a vector kept in registers (the output the next op
    // reads, read by that op), as in_slot and out_slot
    kHotSlot = 0xfe
};

// Output flags per (point, family).
enum : int {
    kFlagZeroSum = 1,     // root vector sums to 0 (gamma_core.cpp:151-152)
    kFlagRecompute64 = 2, // FP32 result not certified: recompute on BATCHED64
    kFlagRecomputePlain = 4  // FP64 value near the subnormal range: recompute on PLAIN64
};

struct PruneParams {
    // work list
    const TileDesc* tiles;
    const int* n_tiles_dev;   // number of tiles (device-resident count)
    int* tile_counter;        // persistent-kernel work counter (zeroed before launch)
    // model
    const BucketDev* buckets;
    const OpDev* ops;
    int n_ops;
    const ChildDev* kids;
    const StepDev* steps;     // the same program as child steps
    int n_steps;
    int n_leaves;
    int n_nodes;
    const int* node_slot;     // [point][n_nodes] matrix slot per branch
    // families
    const int* counts_T;      // [leaf][F] counts, bucket-sorted family order
    int F;                    // number of families
    const int* perm;          // bucket-sorted index -> caller's family index
    // matrices; leaf branches read leaf64 / leaf32 (both c-major), which are
    // the matrices with the leaf error model folded in (without one: mt64 and
    // the c-major FP32 copy)
    const double* mt64;
    const float* mt32;
    const double* leaf64;
    const float* leaf32;
    const float* cm32;        // the FP32 matrices c-major (unfolded): lane kernel columns
    long long slot_elems;     // Np * Np
    int Np;
    int n_dev;                // error-model deviations (0: none); folded into leaf64/leaf32
    // root prior per bucket: prior[bucket.prior_off + j]
    const double* prior64;
    const float* prior32;
    int model;                // CAFE_GPU_MODEL_*
    // outputs [point][F] in the caller's family order
    double* out_val;
    int* out_flag;
    // optional root vectors (FP64) [F][root_stride], one point only
    double* root_out;
    int root_stride;
    // packed recompute tiles index families through this list
    const int* fam_list;
    int vec_exp;              // FP32 partial vectors are kept below 2^vec_exp
    // split-FP16 path: matrix halves (row-major, 32-column blocks in lane
    // order) and the renormalisation target of its partial vectors
    const _Float16* mt16h;
    const _Float16* mt16l;
    int vec_exp16;
    // partial-vector scratch: per workgroup n_slots x 16 x ldv values
    void* scratch;
    long long scratch_per_wg;  // elements
    int ldv;
    int n_slots;
    int slot_tiles;           // MFMA: 16-row tiles per stored partial vector
    int zero;                 // always 0 (an offset the compiler cannot fold)
    int plain_lds;            // PLAIN64: partial vectors in dynamic LDS
    // nonzero band of the FP32 row-major pool: per slot and 16-row tile t,
    // band[slot * band_stride + t] = lo | hi << 16, the k-tiles [lo, hi)
    // outside of which the tile's entries are exactly 0 (null: no band)
    const int* band;
    int band_stride;
};

// One node of the reconstruction program (children before parents).
struct ReconNode {
    int node;       // engine node index
    int leaf_col;   // >= 0: a leaf's column in the count rows
    int kid0;       // first child in the program's kids list
    int n_kids;
    int slot;       // matrix slot of the branch above the node (-1: root)
};

// Launch shape of the MFMA pruning kernel: ntw 16-row output tiles per wave,
// w waves per workgroup, partial vectors in LDS or global scratch, dynamic
// LDS bytes.
struct MfmaShape {
    int ntw;
    int w;
    bool lds;
    size_t shm;
    int per_cu;          // workgroups resident on one CU (registers and LDS)
    int grid_mult = 0;   // workgroups per resident slot (0: the planner's default)
    int fb = 1;          // blocks of 16 families per tile (2: FP32 only)
};

// Dynamic-LDS bookkeeping of the scaled FP32 kernel for n_slots partial
// vectors, w waves and fb blocks of 16 families: per slot a flush bound
// (double), an exponent and a renormalisation shift (int) per family, per-wave
// maxima (float) per family, and one nominal-scale exponent.
__host__ __device__ inline int mfma_meta_bytes(int n_slots, int w, bool scaled, int fb = 1)
{
    if (!scaled) return 0;
    const int fw = 16 * fb;
    const int b = n_slots * fw * 8 + n_slots * w * fw * 4 + 2 * n_slots * fw * 4 + n_slots * 4;
    return (b + 31) & ~31;
}

// Dynamic-LDS bookkeeping of the split-FP16 kernel: the node's per-wave
// maxima (float) and per slot the stored vector's exponent (int) per family.
__host__ __device__ inline int f16_meta_bytes(int n_slots, int w)
{
    const int b = w * 16 * 4 + n_slots * 16 * 4;
    return (b + 31) & ~31;
}

// Dynamic-LDS copy of a tile's leaf counts: n_leaves x 16 fb ints.
__host__ __device__ inline int mfma_count_bytes(int n_leaves, int fb = 1)
{
    return (n_leaves * 16 * fb * 4 + 31) & ~31;
}

}  // namespace cafe
