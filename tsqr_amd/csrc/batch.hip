// batch.hip -- planning, graph capture and the C ABI of the batched TSQR
// (tsqr_batch_* in include/tsqr_amd.h).  The kernels are in batch.hpp and
// are compiled per dtype in batch_f32.hip / batch_f64.hip.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <new>
#include <vector>

#include "../../include/tsqr_amd.h"
#include "batch.hpp"
#include "batch_launch.hpp"

using namespace tsqr;

struct tsqr_batch {
    int dtype = 0, method = 0, count = 0;
    size_t ws_bytes = 0;
    void* dmem = nullptr;  // problems, couples, item lists
    BatchDev dev;
    std::vector<BatchProb> probs;  // host copies (for the statistics)
    // captured graph of one run, for one workspace pointer
    hipStream_t cap = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    void* graph_work = nullptr;
};

namespace {

// Panel problem for an m x n column block (n <= 32) of a matrix.
template <typename T>
bool make_prob(int m, int n, void* A, int lda, void* R, int ldr, int method, size_t* offset, BatchProb* p) {
    const int N = padded_cols(n);
    const int H = 64 * rows_per_lane<T>(N);
    // plain (Householder) problems: leaves of one tile, one wave each
    // (batch_leaf_r / batch_leaf_q), and no per-leaf side blocks; mixed
    // problems: superleaves of super_tiles tiles for the per-superleaf Grams
    const bool plain = method != TSQR_BATCH_MIXED;
    if (!make_tree(m, n, H, (plain ? 1 : super_tiles<T>(N)) * H, &p->d)) return false;
    if (plain) p->d.side_leaves = 0;
    p->w = ws_layout(p->d, sizeof(T));
    p->ws = (int64_t)*offset;
    *offset += align256(p->w.total);
    p->A = A;
    p->R = R;
    p->lda = lda;
    p->ldr = ldr;
    p->N = N;
    p->vec = ((uintptr_t)A % 16 == 0) && (((size_t)lda * sizeof(T)) % 16 == 0);
    p->mixed = method == TSQR_BATCH_MIXED;
    p->pre_gram = 0;
    return true;
}

template <typename T>
int plan(tsqr_batch* b, const int* m, const int* n, void* const* A, const int* lda, void* const* R,
         const int* ldr) {
    std::vector<BatchProb> first, second;
    std::vector<BatchCouple> couples;
    // the workspace starts with the two stages' Householder fallback lists
    // (count, pad, then one int2 per superleaf); sized after planning
    size_t off = 0;
    std::vector<int> owner_second;  // matrix index of each second panel
    for (int i = 0; i < b->count; ++i) {
        if (!A[i] || !R[i] || n[i] < 1 || m[i] < n[i] || lda[i] < m[i] || ldr[i] < n[i]) return TSQR_ERR_ARG;
        if (n[i] > 64) return TSQR_ERR_SHAPE;
        const int n1 = n[i] < 32 ? n[i] : 32;
        BatchProb p{};
        if (!make_prob<T>(m[i], n1, A[i], lda[i], R[i], ldr[i], b->method, &off, &p)) return TSQR_ERR_SHAPE;
        first.push_back(p);
        if (n[i] > 32) {
            const int n2 = n[i] - 32;
            T* A2 = reinterpret_cast<T*>(A[i]) + (size_t)32 * lda[i];
            T* R12 = reinterpret_cast<T*>(R[i]) + (size_t)32 * ldr[i];
            BatchProb q{};
            if (!make_prob<T>(m[i], n2, A2, lda[i], R12 + 32, ldr[i], b->method, &off, &q)) return TSQR_ERR_SHAPE;
            second.push_back(q);
            BatchCouple c{};
            c.Q1 = A[i];
            c.A2 = A2;
            c.R12 = R12;
            c.m = m[i];
            c.n2 = n2;
            c.lda = lda[i];
            c.ldr = ldr[i];
            // row blocks of the coupling products: a full superleaf height
            // (the mixed method's second-panel Grams come from them)
            c.leaf_rows = super_tiles<T>(32) * 64 * rows_per_lane<T>(32);
            c.leaves = (m[i] + c.leaf_rows - 1) / c.leaf_rows;
            c.vec = p.vec;
            c.mixed = (sizeof(T) == 4 && b->method == TSQR_BATCH_MIXED) ? 1 : 0;
            c.ws = (int64_t)off;
            off += couple_ws_bytes(c.leaves);
            couples.push_back(c);
        }
    }
    // Mixed FP32: where the second panel's superleaves are the couple's row
    // blocks, batch_project forms their Gram matrices as it writes the
    // projected rows, and the second stage's Gram launch skips that panel.
    for (size_t k = 0; k < couples.size(); ++k) {
        BatchCouple& c = couples[k];
        BatchProb& q = second[k];
        c.gram = (sizeof(T) == 4 && q.mixed && q.d.lv[0].tile_rows == c.leaf_rows && q.d.lv[0].tiles == c.leaves)
                     ? 1
                     : 0;
        q.pre_gram = c.gram;
    }
    std::vector<BatchProb> all(first);
    all.insert(all.end(), second.begin(), second.end());
    size_t leaves_in[2] = {0, 0};
    for (size_t k = 0; k < all.size(); ++k) leaves_in[k < first.size() ? 0 : 1] += all[k].d.lv[0].tiles;
    const size_t list_bytes[2] = {align256(8 + leaves_in[0] * sizeof(int2)), align256(8 + leaves_in[1] * sizeof(int2))};
    const size_t head = list_bytes[0] + list_bytes[1];
    for (auto& p : all) p.ws += (int64_t)head;
    for (size_t k = 0; k < couples.size(); ++k) {
        BatchCouple& c = couples[k];
        c.ws += (int64_t)head;
        const BatchProb& q = all[first.size() + k];
        c.side2 = q.ws + (int64_t)q.w.side;
    }
    off += head;

    // item lists (problem indices into `all`)
    struct Stage {
        std::vector<int2> probs, leaves, groups, level1;
        std::vector<std::vector<int2>> down, up;
    } st[2];
    for (int k = 0; k < (int)all.size(); ++k) {
        Stage& s = st[k < (int)first.size() ? 0 : 1];
        const TreeDesc& d = all[k].d;
        s.probs.push_back(make_int2(k, 0));
        for (int t = 0; t < d.lv[0].tiles; ++t) s.leaves.push_back(make_int2(k, t));
        for (int g = 0; g < batch_groups(d.lv[0].tiles); ++g) s.groups.push_back(make_int2(k, g));
        if (d.levels >= 2)
            for (int t = 0; t < d.lv[1].tiles; ++t) s.level1.push_back(make_int2(k, t));
        for (int l = 1; l <= d.levels - 2; ++l) {
            if ((int)s.down.size() <= l) s.down.resize(l + 1);
            for (int t = 0; t < d.lv[l].tiles; ++t) s.down[l].push_back(make_int2(k, t));
        }
        for (int l = 1; l <= d.levels - 1; ++l) {
            if ((int)s.up.size() <= l) s.up.resize(l + 1);
            for (int t = 0; t < d.lv[l].tiles; ++t) s.up[l].push_back(make_int2(k, t));
        }
    }
    std::vector<int2> crow, clist;
    for (int c = 0; c < (int)couples.size(); ++c) {
        clist.push_back(make_int2(c, 0));
        for (int g = 0; g < batch_groups(couples[c].leaves); ++g) crow.push_back(make_int2(c, g));
    }

    // one device allocation: problems, couples, then every item list
    size_t bytes = align256(all.size() * sizeof(BatchProb)) + align256(couples.size() * sizeof(BatchCouple));
    size_t nitems = crow.size() + clist.size();
    for (auto& s : st) {
        nitems += s.probs.size() + s.leaves.size() + s.groups.size() + s.level1.size();
        for (auto& d : s.down) nitems += d.size();
        for (auto& u : s.up) nitems += u.size();
    }
    bytes += nitems * sizeof(int2) + 256;
    std::vector<char> host(bytes, 0);
    size_t pos = 0;
    auto put = [&](const void* src, size_t n_bytes) {
        const size_t at = pos;
        if (n_bytes) std::copy((const char*)src, (const char*)src + n_bytes, host.begin() + at);
        pos += n_bytes;
        return at;
    };
    const size_t at_probs = put(all.data(), all.size() * sizeof(BatchProb));
    pos = align256(pos);
    const size_t at_couples = put(couples.data(), couples.size() * sizeof(BatchCouple));
    pos = align256(pos);
    struct At {
        size_t at;
        int count;
    };
    auto put_items = [&](const std::vector<int2>& v) { return At{put(v.data(), v.size() * sizeof(int2)), (int)v.size()}; };
    At a_crow = put_items(crow);
    At a_clist = put_items(clist);
    At a_probs[2], a_leaves[2], a_groups[2], a_l1[2];
    std::vector<At> a_down[2], a_up[2];
    for (int s = 0; s < 2; ++s) {
        a_probs[s] = put_items(st[s].probs);
        a_leaves[s] = put_items(st[s].leaves);
        a_groups[s] = put_items(st[s].groups);
        a_l1[s] = put_items(st[s].level1);
        for (auto& d : st[s].down) a_down[s].push_back(put_items(d));
        for (auto& u : st[s].up) a_up[s].push_back(put_items(u));
    }
    void* dmem = nullptr;
    if (hipMalloc(&dmem, bytes) != hipSuccess) return TSQR_ERR_ALLOC;
    if (hipMemcpy(dmem, host.data(), bytes, hipMemcpyHostToDevice) != hipSuccess) {
        (void)hipFree(dmem);
        return TSQR_ERR_HIP;
    }
    char* base = reinterpret_cast<char*>(dmem);
    auto list = [&](const At& a) { return ItemList{reinterpret_cast<const int2*>(base + a.at), a.count}; };
    BatchDev& dev = b->dev;
    dev.method = b->method;
    dev.probs = reinterpret_cast<const BatchProb*>(base + at_probs);
    dev.nprobs = (int)all.size();
    dev.couples = reinterpret_cast<const BatchCouple*>(base + at_couples);
    dev.ncouples = (int)couples.size();
    dev.couple_rows = list(a_crow);
    dev.couple_list = list(a_clist);
    for (int s = 0; s < 2; ++s) {
        dev.stage[s].fallback = s == 0 ? 0 : (int64_t)list_bytes[0];
        dev.stage[s].probs = list(a_probs[s]);
        dev.stage[s].leaves = list(a_leaves[s]);
        dev.stage[s].groups = list(a_groups[s]);
        dev.stage[s].level1 = list(a_l1[s]);
        dev.stage[s].down.clear();
        for (auto& a : a_down[s]) dev.stage[s].down.push_back(list(a));
        dev.stage[s].up.clear();
        for (auto& a : a_up[s]) dev.stage[s].up.push_back(list(a));
    }
    b->dmem = dmem;
    b->ws_bytes = off ? off : 256;
    b->probs = all;
    return TSQR_OK;
}

void release_graph(tsqr_batch* b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
    b->graph_work = nullptr;
}

int run(tsqr_batch* b, void* work, hipStream_t st) {
    return b->dtype == TSQR_FLOAT ? batch_launch<float>(b->dev, work, st) : batch_launch<double>(b->dev, work, st);
}

}  // namespace

extern "C" {

int tsqr_batch_create(int dtype, int method, int count, const int* m, const int* n, void* const* A, const int* lda,
                      void* const* R, const int* ldr, tsqr_batch** out) {
    if (!out || count < 1 || !m || !n || !A || !lda || !R || !ldr) return TSQR_ERR_ARG;
    if (dtype != TSQR_FLOAT && dtype != TSQR_DOUBLE) return TSQR_ERR_ARG;
    if (method != TSQR_BATCH_PLAIN && method != TSQR_BATCH_MIXED) return TSQR_ERR_ARG;
    if (method == TSQR_BATCH_MIXED && dtype != TSQR_FLOAT) return TSQR_ERR_ARG;
    *out = nullptr;
    tsqr_batch* b = new (std::nothrow) tsqr_batch();
    if (!b) return TSQR_ERR_ALLOC;
    b->dtype = dtype;
    b->method = method;
    b->count = count;
    const int e = dtype == TSQR_FLOAT ? plan<float>(b, m, n, A, lda, R, ldr) : plan<double>(b, m, n, A, lda, R, ldr);
    if (e != TSQR_OK) {
        delete b;
        return e;
    }
    *out = b;
    return TSQR_OK;
}

int tsqr_batch_set_mixed(tsqr_batch* b, double gate, double tol) {
    if (!b || !(gate >= 1.0) || !(tol > 0.0)) return TSQR_ERR_ARG;
    b->dev.gate = gate;
    b->dev.tol = tol;
    release_graph(b);  // the captured graph holds the old values
    return TSQR_OK;
}

size_t tsqr_batch_workspace_bytes(const tsqr_batch* b) { return b ? b->ws_bytes : 0; }

int tsqr_batch_run(tsqr_batch* b, void* work, size_t work_bytes, void* stream) {
    if (!b || !work) return TSQR_ERR_ARG;
    if (work_bytes < b->ws_bytes) return TSQR_ERR_WORKSPACE;
    return run(b, work, (hipStream_t)stream);
}

int tsqr_batch_graph_launch(tsqr_batch* b, void* work, size_t work_bytes, void* stream) {
    if (!b || !work) return TSQR_ERR_ARG;
    if (work_bytes < b->ws_bytes) return TSQR_ERR_WORKSPACE;
    if (!b->exec || b->graph_work != work) {
        release_graph(b);
        if (!b->cap && hipStreamCreateWithFlags(&b->cap, hipStreamNonBlocking) != hipSuccess) return TSQR_ERR_HIP;
        if (hipStreamBeginCapture(b->cap, hipStreamCaptureModeThreadLocal) != hipSuccess) return TSQR_ERR_HIP;
        const int e = run(b, work, b->cap);
        hipGraph_t g = nullptr;
        const hipError_t ce = hipStreamEndCapture(b->cap, &g);
        if (e != TSQR_OK || ce != hipSuccess) {
            if (g) (void)hipGraphDestroy(g);
            return e != TSQR_OK ? e : TSQR_ERR_HIP;
        }
        b->graph = g;
        if (hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0) != hipSuccess) {
            release_graph(b);
            return TSQR_ERR_HIP;
        }
        b->graph_work = work;
    }
    return hipGraphLaunch(b->exec, (hipStream_t)stream) == hipSuccess ? TSQR_OK : TSQR_ERR_HIP;
}

int tsqr_batch_stats(const tsqr_batch* b, const void* work, int* cholesky, int* householder, int* second_pass,
                     int* check_failed) {
    if (!b || !work || !cholesky || !householder || !second_pass || !check_failed) return TSQR_ERR_ARG;
    *cholesky = *householder = *second_pass = *check_failed = 0;
    if (hipDeviceSynchronize() != hipSuccess) return TSQR_ERR_HIP;
    for (const BatchProb& p : b->probs) {
        int ctl[3];
        if (hipMemcpy(ctl, reinterpret_cast<const char*>(work) + p.ws + p.w.ctl, sizeof(ctl),
                      hipMemcpyDeviceToHost) != hipSuccess)
            return TSQR_ERR_HIP;
        if (ctl[0]) {
            *cholesky += 1;
            *second_pass += ctl[1] ? 1 : 0;
            *check_failed += ctl[2] ? 1 : 0;
        } else {
            *householder += 1;
        }
    }
    return TSQR_OK;
}

int tsqr_batch_q_forms(const tsqr_batch* b, const void* work, int* bf16_pass1, int* bf16_pass2) {
    if (!b || !work || !bf16_pass1 || !bf16_pass2) return TSQR_ERR_ARG;
    *bf16_pass1 = *bf16_pass2 = 0;
    if (hipDeviceSynchronize() != hipSuccess) return TSQR_ERR_HIP;
    for (const BatchProb& p : b->probs) {
        int ctl[5];
        if (hipMemcpy(ctl, reinterpret_cast<const char*>(work) + p.ws + p.w.ctl, sizeof(ctl),
                      hipMemcpyDeviceToHost) != hipSuccess)
            return TSQR_ERR_HIP;
        if (p.mixed && ctl[0]) {
            *bf16_pass1 += ctl[kCtlBf16Q1] ? 1 : 0;
            *bf16_pass2 += (ctl[1] && ctl[kCtlBf16Q2]) ? 1 : 0;
        }
    }
    return TSQR_OK;
}

int tsqr_batch_destroy(tsqr_batch* b) {
    if (!b) return TSQR_OK;
    release_graph(b);
    if (b->cap) (void)hipStreamDestroy(b->cap);
    if (b->dmem) (void)hipFree(b->dmem);
    delete b;
    return TSQR_OK;
}

}  // extern "C"
