// plan.hpp -- host-side shape of the TSQR reduction tree and its workspace.
//
// The reference walks a tree whose level heights are H_{l+1} = ceil(H_l/BS)*n
// (kernelQR.h:341-388) inside chunks of NUM_SM*BS rows, and adds a second
// TSQR over the stacked chunk Rs when m is larger than one chunk
// (TallShinnyQR.h:106-127).  Here the tree is one structure over all m rows:
//   level 0   superleaves of leaf_rows rows of A (the last one may be
//             shorter): between 1 and ST tiles of H rows, H = 64*RPL the
//             Householder tile height and ST = min(8, H/N);
//   level l>0 blocks of arity*n rows, arity = floor(H/n); child b of level l
//             contributes min(rows_b, n) rows of its R at row (b % arity)*n
//             of block b / arity on level l+1;
//   root      the first level with a single tile.
// There is no chunking, hence no size cap: the tree just gets deeper.
//
// Blocks of level l >= 1: tiles_l blocks of H x n elements, column-major
// with leading dimension H.  A block first holds the stacked child Rs, then
// its reflectors, then (after the down sweep) the S factors of its children.
#pragma once
#include <stddef.h>
#include <stdint.h>

namespace tsqr {

constexpr int kMaxLevels = 40;

struct LevelDesc {
    int64_t rows;     // valid rows on this level
    int64_t blk_off;  // element offset of this level's blocks (levels >= 1)
    int tiles;        // tiles on this level
    int tile_rows;    // rows per full tile on this level
    int ctr_off;      // offset of this level's tickets (levels >= 1)
    int pad;
};

struct TreeDesc {
    int m, n;
    int H;       // tile height = 64 * RPL = leading dimension of a block
    int arity;   // children per parent
    int levels;  // number of levels (root is levels - 1)
    int total_ctr;
    // FP64 32 x 32 side blocks kept per level-0 superleaf (the superleaf
    // Gram paths), then the whole-matrix blocks; none per leaf in plans with
    // single-tile leaves.
    int side_leaves;
    int64_t block_elems;  // elements of all blocks
    LevelDesc lv[kMaxLevels];
};

inline size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

// Builds the tree for an m x n problem: level-0 tiles of leaf_rows rows,
// blocks of H rows above, each holding up to `arity` child Rs: H / n (a
// full block), or fewer when arity_cap (>= 2) asks for a narrower, deeper
// tree (the autotuner's knob).  Returns false if the shape cannot be
// represented (arity < 2 or too many levels).
inline bool make_tree(int m, int n, int H, int leaf_rows, TreeDesc* d, int arity_cap = 0) {
    d->m = m;
    d->n = n;
    d->H = H;
    d->arity = H / n;
    if (arity_cap >= 2 && arity_cap < d->arity) d->arity = arity_cap;
    if (d->arity < 2) return false;
    int l = 0;
    int64_t rows = m;
    int tile_rows = leaf_rows;
    int64_t blk = 0;
    int ctr = 0;
    for (;;) {
        if (l >= kMaxLevels) return false;
        LevelDesc& L = d->lv[l];
        L.rows = rows;
        L.tile_rows = tile_rows;
        L.tiles = (int)((rows + tile_rows - 1) / tile_rows);
        L.pad = 0;
        if (l == 0) {
            L.blk_off = 0;
            L.ctr_off = 0;
        } else {
            L.blk_off = blk;
            blk += (int64_t)L.tiles * H * n;
            L.ctr_off = ctr;
            ctr += L.tiles;
        }
        if (L.tiles == 1) break;
        const int64_t last = rows - (int64_t)(L.tiles - 1) * tile_rows;
        rows = (int64_t)(L.tiles - 1) * n + (last < n ? last : n);
        tile_rows = d->arity * n;
        ++l;
    }
    d->levels = l + 1;
    d->total_ctr = ctr;
    d->side_leaves = d->lv[0].tiles;
    d->block_elems = blk;
    return true;
}

// Workspace layout of one TSQR call (byte offsets, 256-byte aligned):
//   blocks   tree blocks of levels >= 1 (elements of T)
//   tickets  one int per tile of levels >= 1
//   flags    one int per level-0 superleaf: 0 Gram path, 1 Householder path
//   list     count, then the indices of the Householder-path superleaves
//   ctl      control words: ctl[0] = 1 when the whole matrix took the flat
//            Gram path; ctl[1] .. ctl[4] for the mixed path (mixed.hpp);
//            ctl[64 .. 128) .. ctl[256 .. 320) the tickets of four streaming
//            passes (the Gram passes and the mixed path's Q passes)
//   partial  32 x 32 FP64 partial sums: min(superleaves, kGramPartials)
//            blocks of superleaf Gram sums plus kGramPartials2 for a second
//            stage when there are more than kGramPartials2 of them, or the
//            streaming Gram kernel's P + ceil(P / 32) blocks (P =
//            stream_partials(m)), whichever is more
//   side     32 x 32 FP64 per superleaf (the Gram matrix, then R^{-1}; none
//            for single-tile leaves: side_leaves), then three more: the
//            whole matrix's R^{-1} on the flat path (T0 = side_leaves), its
//            R (T0+1) and the mixed path's residuals (T0+2)
constexpr int kGramPartials = 1024;
constexpr int kGramPartials2 = 32;
constexpr int kCtlWords = 320;          // control words, including the stream tickets
constexpr int kCtlStreamTickets = 64;   // ctl[64 ..]: tickets of the streaming Gram
constexpr int kCtlStreamTickets2 = 128; // ctl[128 ..]: tickets of a second streaming pass
constexpr int kCtlStreamTickets3 = 192; // ctl[192 ..], ctl[256 ..]: tickets of the mixed path's
constexpr int kCtlStreamTickets4 = 256; //   two streamed Q passes
// ctl[3], ctl[4]: the FP32 mixed path's Q pass 1 / pass 2 may run on the BF16
// matrix cores (set where R^{-1} is well conditioned: kappa_est <= 2n after
// the first Cholesky, ||Q1^T Q1 - I||_F <= kBf16QMaxResidual before the
// second)
constexpr int kCtlBf16Q1 = 3;
constexpr int kCtlBf16Q2 = 4;
constexpr double kBf16QMaxResidual = 0.5;
// Gate of the FP32 mixed path's FP64-Gram escalation (taken where the
// BF16x3 Gram's gate rejects): the FP64 Gram of FP32 data has exact
// products, so R1 is accurate to about u64 kappa^2 relative, and Q1 = A
// R1^{-1} is orthonormal to O(u64 kappa_est^2) <= 1e-2 below the gate,
// which the second CholeskyQR pass restores to FP32 level.  A caller's gate
// below the default kMixedGate (tsqr_set_mixed, tsqr_set_leaf_path) limits
// this pass too, so a gate of 1 still forces the Householder fallback.
constexpr double kMixedGate = 1024.0;
constexpr double kMixedGram64Gate = 1e7;
inline double mixed_gram64_gate(double gate) {
    return gate < kMixedGate ? gate : (gate > kMixedGram64Gate ? gate : kMixedGram64Gate);
}

// The streaming Gram kernel (tsqr_gram_stream, gram.hpp) runs P workgroups
// of 4 waves, P = stream_groups(m): at most kStreamMaxGroups, and four
// 64-row chunks per wave, so that the loads of one chunk overlap the MFMAs
// of the previous one, unless that leaves CUs idle (fewer than
// kStreamFillGroups workgroups): then down to one chunk per wave.  Its
// partial sums take P blocks plus one per group (stream_group_size).
constexpr int kStreamMaxGroups = 1024;
constexpr int kStreamFillGroups = 256;
constexpr int kStreamGroup = 32;  // the largest group
static_assert(kStreamMaxGroups / kStreamGroup + 1 <= kCtlStreamTickets, "stream tickets do not fit");

// Workgroups per group of the streaming Gram's partial sums for a grid of P:
// small grids use small groups, so that the last workgroup of a group and
// the final sum each add only a few partials (at most 32 groups either way).
constexpr int stream_group_size(int P) { return P <= 64 ? 8 : (P <= 256 ? 16 : kStreamGroup); }
constexpr int stream_group_count(int P) { return (P + stream_group_size(P) - 1) / stream_group_size(P); }
static_assert(stream_group_count(64) <= 32 && stream_group_count(256) <= 32 &&
                  stream_group_count(kStreamMaxGroups) <= 32,
              "at most 32 stream groups");

// Largest grid of the streaming passes, kStreamMaxGroups unless tuned lower
// (tsqr_set_stream_shape; defined in tsqr_amd.hip).
int tuned_stream_max_groups();

inline int stream_groups(int64_t m) {
    const int64_t chunks = (m + 63) / 64;
    int64_t p = (chunks + 15) / 16;
    if (p < kStreamFillGroups) {
        const int64_t one = (chunks + 3) / 4;  // one chunk per wave
        p = one < kStreamFillGroups ? one : kStreamFillGroups;
    }
    if (p < 1) p = 1;
    const int cap = tuned_stream_max_groups();
    return (int)(p < cap ? p : cap);
}

// Short matrices: the R pass without tickets, one chunk of chunk_rows rows
// (64 FP32, 32 FP64) per wave, on workgroups of the fewest waves (4, 8 or
// 16) that keep the grid within kShortGroups (tsqr_gram_short); their
// partials tsqr_flat_short sums with 8 waves.  A ticketed launch spends
// about 8 us of a 27648 x 32 call on the release fence, the ticket and the
// group sums; fewer, wider workgroups leave fewer partials but put the FP64
// products on fewer CUs.  short_waves is 0 above 16 kShortGroups chunks.
constexpr int kShortGroups = 64;
inline int short_waves(int64_t m, int chunk_rows) {
    const int64_t chunks = (m + chunk_rows - 1) / chunk_rows;
    for (int w = 4; w <= 16; w *= 2)
        if ((chunks + w - 1) / w <= kShortGroups) return w;
    return 0;
}
inline int short_groups(int64_t m, int chunk_rows) {
    const int w = short_waves(m, chunk_rows);
    const int64_t chunks = (m + chunk_rows - 1) / chunk_rows;
    return w ? (int)((chunks + w - 1) / w) : 0;
}

// The FP32 flat path's passes on 4 x 4 tiles (stream4.hpp).  The R pass
// (tsqr_gram_stream4) takes over from tsqr_gram_stream at kStream4MinRows
// rows: below that its longer reduction at the end cost more than its
// cheaper products saved (27648 x 32: 23 against 14 us; 1M x 32: 56
// against 70 us).  Its grid doubles the cap (two workgroups per CU) once
// every wave has at least 64 chunks of 64 rows (23.9M x 32: 0.515 against
// 0.552 ms; 4.2M x 32: 0.131 against 0.121 ms, so not there).  The last
// kCachedChunks chunks of its sweep are read with plain loads, so that they
// stay in the memory-side cache (256 MB on MI355X) for the Q pass.
constexpr int64_t kStream4MinRows = (int64_t)1 << 19;
constexpr int kCachedChunks = 24576;  // 192 MB of 32 FP32 columns
inline bool use_stream4(int64_t m) { return m >= kStream4MinRows; }
inline int stream_groups4(int64_t m) {
    const int64_t chunks = m / 64;
    const int cap = tuned_stream_max_groups();
    const int big = 2 * cap < kStreamMaxGroups ? 2 * cap : kStreamMaxGroups;
    return chunks >= (int64_t)64 * 4 * big ? big : stream_groups(m);
}
// Grid of the FP32 flat Q pass (tsqr_flat_q4): one chunk per wave up to the
// cap, or up to twice the cap below 2048 chunks (measured: 100000 x 32 runs
// fastest with 391 workgroups, 262144 x 32 and taller with 256).
inline int flat_q_groups(int64_t m) {
    const int64_t chunks = (m + 63) / 64;
    const int cap = tuned_stream_max_groups();
    int64_t lim = chunks < 2048 ? 2 * (int64_t)cap : cap;
    if (lim > kStreamMaxGroups) lim = kStreamMaxGroups;
    int64_t p = (chunks + 3) / 4;
    if (p > lim) p = lim;
    return (int)(p < 1 ? 1 : p);
}

// Partial blocks of a streaming Gram launch of P workgroups: one per
// workgroup plus one per group; the workspace holds enough for either grid.
inline size_t stream_partials(int64_t m) {
    const int p = stream_groups(m);
    const int p4 = stream_groups4(m);
    const size_t a = (size_t)(p + stream_group_count(p));
    const size_t b = (size_t)(p4 + stream_group_count(p4));
    return a > b ? a : b;
}

// Workspace of the device-side checker (check.hpp): three partial areas of
// stream_partials(m) blocks and three sets of stream tickets.
inline size_t check_workspace_bytes(int64_t m) {
    return 3 * stream_partials(m) * 1024 * sizeof(double) + 3 * kCtlStreamTickets * sizeof(int);
}

// The set-up of one TSQR call's workspace: tickets = 0, flags = value, the
// Householder list empty (value 0) or holding every superleaf (value 1), the
// control words cleared.  Run by init_flags, or folded into the first kernel
// of a call (tsqr_gram_short) where that kernel uses none of these words.
struct WsInit {
    int* tickets;
    int nt;
    int* flags;
    int* list;
    int nf;
    int value;
    int* ctl;
    bool on;  // folded into the kernel that receives it
    __host__ __device__ int count() const {
        const int a = nt > nf ? nt : nf;
        return a > kCtlWords ? a : kCtlWords;
    }
    __device__ inline void at(int i) const {
        if (i < nt) tickets[i] = 0;
        if (i < nf) {
            flags[i] = value;
            if (value) list[1 + i] = i;
        }
        if (i < kCtlWords) ctl[i] = 0;
        if (i == 0) list[0] = value ? nf : 0;
    }
};

struct WsLayout {
    size_t blocks, tickets, flags, list, ctl, partial, side, total;
};

inline WsLayout ws_layout(const TreeDesc& d, size_t elem) {
    WsLayout w;
    const size_t leaves = (size_t)d.lv[0].tiles;
    w.blocks = 0;
    w.tickets = align256((size_t)d.block_elems * elem);
    w.flags = w.tickets + align256((size_t)d.total_ctr * sizeof(int));
    w.list = w.flags + align256(leaves * sizeof(int));
    w.ctl = w.list + align256((leaves + 1) * sizeof(int));
    w.partial = w.ctl + kCtlWords * sizeof(int);
    const size_t sl = (size_t)d.side_leaves;
    const size_t parts = sl < (size_t)kGramPartials ? sl : (size_t)kGramPartials;
    const size_t parts2 = parts > (size_t)kGramPartials2 ? (size_t)kGramPartials2 : 0;
    // the streaming passes' partials, or the short R pass's (chunks of 64
    // FP32 or 32 FP64 rows)
    size_t sparts = stream_partials(d.m);
    const size_t shorts = (size_t)short_groups(d.m, elem == 4 ? 64 : 32);
    if (shorts > sparts) sparts = shorts;
    const size_t nparts = parts + parts2 > sparts ? parts + parts2 : sparts;
    w.side = w.partial + nparts * 1024 * sizeof(double);
    w.total = w.side + align256((sl + 3) * 1024 * sizeof(double));
    return w;
}

inline size_t tree_workspace_bytes(const TreeDesc& d, size_t elem) { return ws_layout(d, elem).total; }

}  // namespace tsqr
