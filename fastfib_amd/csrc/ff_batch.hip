// Batched products: many independent products of mixed sizes, as the
// reference's callers would issue them one big_mul at a time.
//
// FF_BATCH_PLAIN runs them one after another, each as its own launch
// sequence (forward passes, inner convolution, inverse passes, CRT). The
// grouped forms build, once, a work list per pass kind: every transform
// array (or product) of the batch with its own schedule, the flat grid cut
// among them by a start array, largest first; one launch then covers the
// whole batch for that pass kind (about a dozen launches per batch however
// many products). FF_BATCH_GRAPH captures those launches in a HIP graph and
// replays it. Schedules come from the reference rule, the autotuner table
// or the learned dispatcher (ff_tune.hip).
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <set>
#include <tuple>
#include <vector>

#include "../../include/fastfib_amd.h"
#include "ff_engine.h"
#include "ff_internal.h"
#include "ff_limbs.h"
#include "ff_runtime.h"

namespace {

struct BLaunch {
    int kind;  // 0 outer list, 1 inner list, 2 crt list, 3 zero fill
    bool inv, top, r3, small;
    int sh;    // outer list: the items' fixed-geometry shape (ff_outer_shape), or 0;
               // inner list: 1 where every block is 2^12 words (the fixed-path kernel)
    const void *items;
    const uint32_t *start;
    int n;
    uint32_t blocks;
    size_t lds;
    double bytes, conv_total, out_total;
    uint64_t *fill;
};

size_t al(size_t b) { return (b + 255) / 256 * 256; }

}  // namespace

struct ff_batch {
    int flags = 0;
    std::vector<ff_prod> items;
    std::vector<ProdShape> shape;
    std::vector<FFSched> sched;
    std::vector<char> zero;
    std::vector<char> school;  // both operands below the schoolbook bound: no transforms
    std::vector<void *> work;
    unsigned long long *slots = nullptr;
    char *desc = nullptr;
    std::vector<BLaunch> launches;
    std::vector<MulItem> plain;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    // the work-queue form
    const FFQTask *qtasks = nullptr;  // in desc
    uint32_t ntasks = 0, nctr = 0;
    uint32_t *ctr = nullptr;
    size_t qlds = 0;
    double qbytes = 0;
};

// Bit lengths of many device numbers: thread k scans number k down from its
// top limb.
__global__ void k_bits(const uint64_t *const *ptr, const uint64_t *len, int n, uint64_t *bits) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const uint64_t *d = ptr[k];
    uint64_t i = len[k];
    while (i > 0 && d[i - 1] == 0) i--;
    bits[k] = i ? 64 * (i - 1) + (64 - __clzll((long long)d[i - 1])) : 0;
}

static void read_bits(std::vector<ff_prod> &items) {
    std::vector<const uint64_t *> ptr;
    std::vector<uint64_t> len;
    std::vector<std::pair<size_t, int>> who;
    for (size_t i = 0; i < items.size(); i++) {
        if (!items[i].bits_a) ptr.push_back(items[i].a), len.push_back(items[i].na), who.push_back({i, 0});
        const bool sq = !items[i].b || items[i].b == items[i].a;
        if (!sq && !items[i].bits_b) ptr.push_back(items[i].b), len.push_back(items[i].nb), who.push_back({i, 1});
    }
    if (ptr.empty()) return;
    const int n = (int)ptr.size();
    char *buf = (char *)ff_dalloc(al(n * 8) * 3);
    const uint64_t **dptr = (const uint64_t **)buf;
    uint64_t *dlen = (uint64_t *)(buf + al(n * 8)), *dbits = (uint64_t *)(buf + 2 * al(n * 8));
    FF_CHECK(hipMemcpyAsync(dptr, ptr.data(), n * 8, hipMemcpyHostToDevice, g_st));
    FF_CHECK(hipMemcpyAsync(dlen, len.data(), n * 8, hipMemcpyHostToDevice, g_st));
    hipLaunchKernelGGL(k_bits, dim3((n + 255) / 256), dim3(256), 0, g_st, dptr, dlen, n, dbits);
    FF_CHECK(hipGetLastError());
    std::vector<uint64_t> bits(n);
    FF_CHECK(hipMemcpyAsync(bits.data(), dbits, n * 8, hipMemcpyDeviceToHost, g_st));
    FF_CHECK(hipStreamSynchronize(g_st));
    ff_dfree(buf);
    for (int k = 0; k < n; k++) {
        if (who[k].second == 0) items[who[k].first].bits_a = bits[k];
        else items[who[k].first].bits_b = bits[k];
    }
}

// Lays out the work lists of a grouped batch and uploads them in one copy.
static void build_lists(ff_batch *B) {
    const size_t n = B->items.size();
    std::vector<size_t> order;
    for (size_t i = 0; i < n; i++)
        if (!B->zero[i] && !B->school[i]) order.push_back(i);
    std::sort(order.begin(), order.end(), [&](size_t x, size_t y) { return B->shape[x].L > B->shape[y].L; });

    std::map<std::tuple<uint64_t, unsigned, int, int>, FFPlan> plans;
    std::vector<const FFPlan *> plan(n, nullptr);
    std::vector<uint32_t *> fa(n), fb(n);
    std::vector<char> own_b(n, 0);  // the item transforms its b (else it reads an earlier item's)
    std::vector<int> b_owner(n, -1);  // the item whose b transform this one reads
    std::vector<int> pre_of(n, -1);   // the shared one-block spectrum this one reads
    std::map<const uint32_t *, int> owner_of;
    // products sharing their b operand (same pointer, limbs, width and
    // schedule) share one transform of it
    std::map<std::tuple<const uint64_t *, uint64_t, uint64_t, unsigned, const FFPlan *>, uint32_t *> shared_b;
    std::vector<FFCrtJob> jobs(n);
    ff_lb_begin();
    // every item's transform arrays come from one block (one allocation per
    // batch, not one per item: a decimal level has tens of thousands of
    // items): the first walk sizes it, the second below carves it
    size_t arena_bytes = 0;
    {
        std::set<std::tuple<const uint64_t *, uint64_t, uint64_t, unsigned, uint64_t, int, int>> seen_b;
        for (size_t i : order) {
            const ProdShape &s = B->shape[i];
            const ff_prod &p = B->items[i];
            const bool sq = !p.b || p.b == p.a;
            const size_t words = 3 * (size_t)s.L;
            bool ownb = false;
            if (!sq) {
                auto bk = std::make_tuple(p.b, (uint64_t)p.nb, (uint64_t)p.bits_b, s.bw, s.L,
                                          ff_sched_inner_log(B->sched[i]), ff_sched_tile_words(B->sched[i]));
                ownb = seen_b.insert(bk).second;
            }
            arena_bytes += al(words * 4) + (ownb ? al(words * 4) : 0) + 256;
        }
    }
    char *arena = (char *)ff_dalloc(arena_bytes ? arena_bytes : 256);
    B->work.assign(1, arena);
    size_t arena_off = 0;
    for (size_t i : order) {
        const ProdShape &s = B->shape[i];
        const int il = ff_sched_inner_log(B->sched[i]), tw = ff_sched_tile_words(B->sched[i]);
        auto key = std::make_tuple(s.L, s.bw, il, tw);
        auto it = plans.find(key);
        if (it == plans.end()) {
            FFPlan pl;
            if (ff_plan_build(&pl, (uint32_t)s.L, (int)s.bw, s.lg, s.r3, s.conv, il, tw) != 0) {
                fprintf(stderr, "fastfib_amd: no schedule for L=%llu\n", (unsigned long long)s.L);
                exit(1);
            }
            it = plans.emplace(key, pl).first;
        }
        plan[i] = &it->second;
        const ff_prod &p = B->items[i];
        const bool sq = !p.b || p.b == p.a;
        const size_t words = 3 * (size_t)s.L;
        const uint64_t nout = p.na + (sq ? p.na : p.nb) + 1;
        uint32_t *sb = nullptr;
        if (!sq) {
            auto bk = std::make_tuple(p.b, (uint64_t)p.nb, (uint64_t)p.bits_b, s.bw, plan[i]);
            auto f = shared_b.find(bk);
            if (f != shared_b.end()) sb = f->second;
        }
        const bool ownb = !sq && !sb;
        const size_t o_fb = al(words * 4);
        char *w = arena + arena_off;
        arena_off += o_fb + (ownb ? al(words * 4) : 0) + 256;
        if (arena_off > arena_bytes) {
            fprintf(stderr, "fastfib_amd: batch arena overrun\n");
            exit(1);
        }
        fa[i] = (uint32_t *)w;
        fb[i] = sq ? nullptr : (ownb ? (uint32_t *)(w + o_fb) : sb);
        own_b[i] = ownb;
        if (ownb) owner_of[fb[i]] = (int)i;
        else if (sb) b_owner[i] = owner_of[sb];
        if (ownb) shared_b[std::make_tuple(p.b, (uint64_t)p.nb, (uint64_t)p.bits_b, s.bw, plan[i])] = fb[i];
        jobs[i] = FFCrtJob{fa[i], p.out, ff_lb_take(ff_crt_lb_words(nout)), B->slots + i, nout, s.conv,
                           (uint32_t)s.L, (int)s.bw};
    }

    // one-block products sharing their b operand: b's spectrum once, by a
    // forward-only item launched ahead of them, read by every sharer
    std::vector<FFInnerItem> pre;
    std::vector<const FFPlan *> pre_plan;
    std::vector<int> bspec(n, 0);
    {
        std::map<std::tuple<const uint64_t *, uint64_t, uint64_t, unsigned, const FFPlan *>, std::vector<size_t>> users;
        for (size_t i : order) {
            const ff_prod &p = B->items[i];
            if (!p.b || p.b == p.a || !plan[i]->inner.small) continue;
            users[std::make_tuple(p.b, (uint64_t)p.nb, (uint64_t)p.bits_b, B->shape[i].bw, plan[i])].push_back(i);
        }
        for (auto &u : users) {
            if (u.second.size() < 2) continue;
            const FFPlan &pl = *plan[u.second[0]];
            const ff_prod &p = B->items[u.second[0]];
            uint32_t *spec = (uint32_t *)ff_dalloc(al(3 * (size_t)pl.L * 4));
            B->work.push_back(spec);
            FFInnerItem ii;
            ii.in = pl.inner;
            ii.c = FFConv{(const uint32_t *)p.b, nullptr, (uint32_t)(2 * p.nb), 0, spec, nullptr};
            ii.c.bspec = 2;
            pre.push_back(ii);
            pre_plan.push_back(&pl);
            for (size_t i : u.second) fb[i] = spec, bspec[i] = 1, b_owner[i] = -1, pre_of[i] = (int)pre.size() - 1;
        }
    }

    // host images of every list, then one device buffer
    std::vector<char> img;
    auto put = [&](const void *src, size_t bytes) {
        const size_t off = img.size();
        img.resize(off + al(bytes));
        memcpy(img.data() + off, src, bytes);
        return off;
    };
    struct Pending {
        BLaunch L;
        size_t items_off, start_off;
        size_t item_size;
        std::vector<uint32_t> start;  // host copy: item k owns blocks [start[k], start[k + 1])
        std::vector<int> who;         // each item's product (n + k: shared spectrum k; -1: none)
    };
    std::vector<Pending> pend;

    // A launch allocates its largest item's LDS for every workgroup, so
    // items are grouped by how many workgroups their LDS lets a CU hold;
    // outer passes also by their fixed-geometry shape (ff_outer_shape).
    auto occ = [](size_t lds) -> int {
        const size_t k = lds ? (size_t)163840 / lds : 8;
        return (int)(k < 8 ? k : 8);
    };
    auto outer_occs = [&](int pass, int r3sel) {
        std::set<std::pair<int, int>> out;
        for (size_t i : order) {
            const FFPlan &pl = *plan[i];
            if (pl.npass <= pass) continue;
            const FFPass &ps = pl.pass[pass];
            if (pass == 0 && (int)ps.r3 != r3sel) continue;
            out.insert({occ(ff_outer_lds_bytes(&ps)), ff_outer_shape(&ps, pass == 0)});
        }
        return out;
    };
    auto outer_list = [&](int pass, bool inv, int r3sel, int oc, int sh) {
        std::vector<FFOuterItem> its;
        std::vector<uint32_t> start;
        std::vector<int> who;
        uint32_t blocks = 0;
        size_t lds = 0;
        double bytes = 0;
        for (size_t i : order) {
            const FFPlan &pl = *plan[i];
            if (pl.npass <= pass) continue;
            const FFPass &ps = pl.pass[pass];
            if (pass == 0 && (int)ps.r3 != r3sel) continue;
            if (occ(ff_outer_lds_bytes(&ps)) != oc || ff_outer_shape(&ps, pass == 0) != sh) continue;
            const ff_prod &p = B->items[i];
            const bool sq = !p.b || p.b == p.a;
            for (int op = 0; op < (sq ? 1 : 2); op++) {
                if (op == 1 && !inv && !own_b[i]) continue;  // b transformed once for its sharers
                FFOuterItem oi;
                oi.ps = ps;
                oi.x = FFXform{nullptr, 0, op ? fb[i] : fa[i]};
                if (!inv && pass == 0) {
                    oi.x.src = (const uint32_t *)(op ? p.b : p.a);
                    oi.x.nw = (uint32_t)(2 * (op ? p.nb : p.na));
                    bytes += 3.0 * 8.0 * (op ? p.nb : p.na);
                }
                bytes += (!inv && pass == 0) ? 3.0 * 4.0 * pl.L : 3.0 * 8.0 * pl.L;
                its.push_back(oi);
                start.push_back(blocks);
                who.push_back((int)i);
                blocks += 3 * ps.nblocks * ps.tpb;
                const size_t l = ff_outer_lds_bytes(&ps);
                lds = l > lds ? l : lds;
                if (inv) break;  // the inverse runs on the product's array only
            }
        }
        if (its.empty()) return;
        Pending pd{};
        pd.L.kind = 0, pd.L.inv = inv, pd.L.top = pass == 0, pd.L.r3 = r3sel, pd.L.sh = sh, pd.L.n = (int)its.size();
        pd.L.blocks = blocks, pd.L.lds = lds, pd.L.bytes = bytes;
        pd.items_off = put(its.data(), its.size() * sizeof(FFOuterItem));
        pd.start_off = put(start.data(), start.size() * 4);
        pd.item_size = sizeof(FFOuterItem), pd.start = start, pd.who = who;
        pend.push_back(pd);
    };
    // multi-pass items' inner lists also by whether their blocks are 2^12
    // words (the fixed-path kernel)
    auto fix12 = [](const FFPlan &pl) { return !pl.inner.small && pl.inner.s == 12 ? 1 : 0; };
    auto inner_occs = [&](int small, int r3sel) {
        std::set<std::pair<int, int>> out;
        for (size_t i : order) {
            const FFPlan &pl = *plan[i];
            if (pl.inner.small != small || (small && pl.inner.r3 != r3sel)) continue;
            out.insert({occ(ff_inner_lds_bytes(&pl.inner)), fix12(pl)});
        }
        return out;
    };
    auto inner_list = [&](int small, int r3sel, int oc, int fx) {
        std::vector<FFInnerItem> its;
        std::vector<uint32_t> start;
        std::vector<int> who;
        uint32_t blocks = 0;
        size_t lds = 0;
        double bytes = 0;
        for (size_t i : order) {
            const FFPlan &pl = *plan[i];
            if (pl.inner.small != small || (small && pl.inner.r3 != r3sel)) continue;
            if (occ(ff_inner_lds_bytes(&pl.inner)) != oc || fix12(pl) != fx) continue;
            const ff_prod &p = B->items[i];
            const bool sq = !p.b || p.b == p.a;
            FFInnerItem ii;
            ii.in = pl.inner;
            ii.c = FFConv{(const uint32_t *)p.a, sq ? nullptr : (const uint32_t *)p.b, (uint32_t)(2 * p.na),
                          (uint32_t)(2 * (sq ? p.na : p.nb)), fa[i], fb[i]};
            ii.c.bspec = bspec[i];
            its.push_back(ii);
            start.push_back(blocks);
            who.push_back((int)i);
            blocks += 3 * (small ? 1 : pl.L / pl.inner.S);
            const size_t l = ff_inner_lds_bytes(&pl.inner);
            lds = l > lds ? l : lds;
            bytes += small ? 3.0 * (8.0 * (p.na + (sq ? 0 : p.nb)) + 4.0 * pl.L)
                           : 3.0 * 4.0 * pl.L * (sq ? 2 : 3);
        }
        if (its.empty()) return;
        Pending pd{};
        pd.L.kind = 1, pd.L.small = small, pd.L.r3 = r3sel, pd.L.sh = fx, pd.L.n = (int)its.size();
        pd.L.blocks = blocks, pd.L.lds = lds, pd.L.bytes = bytes;
        pd.items_off = put(its.data(), its.size() * sizeof(FFInnerItem));
        pd.start_off = put(start.data(), start.size() * 4);
        pd.item_size = sizeof(FFInnerItem), pd.start = start, pd.who = who;
        pend.push_back(pd);
    };

    // the shared one-block spectra, by radix and occupancy class
    auto pre_all = [&](int r3sel) {
        std::set<int> occs;
        for (size_t k = 0; k < pre.size(); k++)
            if (pre_plan[k]->inner.r3 == r3sel) occs.insert(occ(ff_inner_lds_bytes(&pre_plan[k]->inner)));
        for (int oc : occs) {
            std::vector<FFInnerItem> its;
            std::vector<uint32_t> start;
            std::vector<int> who;
            uint32_t blocks = 0;
            size_t lds = 0;
            double bytes = 0;
            for (size_t k = 0; k < pre.size(); k++) {
                const FFPlan &pl = *pre_plan[k];
                if (pl.inner.r3 != r3sel || occ(ff_inner_lds_bytes(&pl.inner)) != oc) continue;
                its.push_back(pre[k]);
                start.push_back(blocks);
                who.push_back((int)(n + k));
                blocks += 3;
                const size_t l = ff_inner_lds_bytes(&pl.inner);
                lds = l > lds ? l : lds;
                bytes += 3.0 * (4.0 * pre[k].c.anw + 4.0 * pl.L);
            }
            Pending pd{};
            pd.L.kind = 1, pd.L.small = 1, pd.L.r3 = r3sel, pd.L.n = (int)its.size();
            pd.L.blocks = blocks, pd.L.lds = lds, pd.L.bytes = bytes;
            pd.items_off = put(its.data(), its.size() * sizeof(FFInnerItem));
            pd.start_off = put(start.data(), start.size() * 4);
            pd.item_size = sizeof(FFInnerItem), pd.start = start, pd.who = who;
            pend.push_back(pd);
        }
    };

    // every (pass, radix-3) list once per occupancy class
    auto outer_all = [&](int pass, bool inv, int r3sel) {
        for (const auto &os : outer_occs(pass, r3sel)) outer_list(pass, inv, r3sel, os.first, os.second);
    };
    auto inner_all = [&](int small, int r3sel) {
        for (const auto &os : inner_occs(small, r3sel)) inner_list(small, r3sel, os.first, os.second);
    };
    for (int pass = 0; pass < 3; pass++) {
        if (pass == 0) {
            outer_all(0, false, 1);
            outer_all(0, false, 0);
        } else {
            outer_all(pass, false, 0);
        }
    }
    pre_all(1);
    pre_all(0);
    inner_all(0, 0);
    inner_all(1, 1);
    inner_all(1, 0);
    for (int pass = 2; pass >= 0; pass--) {
        if (pass == 0) {
            outer_all(0, true, 1);
            outer_all(0, true, 0);
        } else {
            outer_all(pass, true, 0);
        }
    }
    {
        std::vector<FFCrtJob> js;
        std::vector<uint32_t> start;
        std::vector<int> who;
        uint32_t blocks = 0;
        double conv = 0, out = 0;
        for (size_t i : order) {
            js.push_back(jobs[i]);
            start.push_back(blocks);
            who.push_back((int)i);
            blocks += (uint32_t)ff_crt_blocks(jobs[i].nout, jobs[i].bw);
            conv += (double)jobs[i].conv;
            out += (double)jobs[i].nout;
        }
        if (!js.empty()) {
            Pending pd{};
            pd.L.kind = 2, pd.L.n = (int)js.size(), pd.L.blocks = blocks;
            pd.L.conv_total = conv, pd.L.out_total = out;
            pd.L.bytes = 12.0 * conv + 8.0 * out;
            pd.items_off = put(js.data(), js.size() * sizeof(FFCrtJob));
            pd.start_off = put(start.data(), start.size() * 4);
            pd.item_size = sizeof(FFCrtJob), pd.start = start, pd.who = who;
            pend.push_back(pd);
        }
    }
    {
        // schoolbook items: one wavefront each up to 128 output limbs, else
        // one workgroup each
        std::vector<FFSchoolItem> sch[2];
        double sbytes[2] = {0, 0};
        for (size_t i = 0; i < n; i++) {
            if (!B->school[i]) continue;
            const ff_prod &p = B->items[i];
            const bool sq = !p.b || p.b == p.a;
            const uint64_t nb = sq ? p.na : p.nb;
            const int wave = p.na + nb + 1 <= FF_SCHOOL_WAVE_OUT;
            sch[wave].push_back(FFSchoolItem{p.a, sq ? p.a : p.b, (uint32_t)p.na, (uint32_t)nb, p.out, B->slots + i});
            sbytes[wave] += 8.0 * (2 * p.na + 2 * nb + 1);
        }
        for (int wave = 0; wave < 2; wave++) {
            if (sch[wave].empty()) continue;
            Pending pd{};
            pd.L.kind = wave ? 5 : 4, pd.L.n = (int)sch[wave].size(), pd.L.bytes = sbytes[wave];
            pd.items_off = put(sch[wave].data(), sch[wave].size() * sizeof(FFSchoolItem));
            pd.start_off = 0;
            pd.item_size = sizeof(FFSchoolItem);
            pend.push_back(pd);
        }
    }
    // the work-queue form: every block of every launch above as a task, in
    // launch order, with its dependencies (k_queue)
    size_t tasks_off = 0;
    if (B->flags & FF_BATCH_QUEUE) {
        const size_t ncnt = n + pre.size();
        std::vector<uint32_t> count(ncnt, 0), fwd_total(ncnt, 0);
        for (const Pending &pd : pend) {
            if (pd.L.kind != 0 || pd.L.inv) continue;
            for (size_t k = 0; k < pd.who.size(); k++) {
                const uint32_t end = k + 1 < pd.start.size() ? pd.start[k + 1] : pd.L.blocks;
                fwd_total[pd.who[k]] += end - pd.start[k];
            }
        }
        std::vector<FFQTask> tasks;
        size_t qlds = 0;
        double qbytes = 0;
        for (const Pending &pd : pend) {
            qbytes += pd.L.bytes;
            const int kind = pd.L.kind;
            if (kind == 4 || kind == 5) {
                qlds = std::max(qlds, kind == 4 ? ff_queue_lds_school() : ff_queue_lds_school_wave());
                const uint32_t per = kind == 4 ? 1 : 4;
                for (uint32_t k = 0; k < (uint32_t)pd.L.n; k += per) {
                    const uint32_t m = std::min(per, (uint32_t)pd.L.n - k);
                    tasks.push_back(FFQTask{kind == 4 ? (uint32_t)FQ_SCHOOL : (uint32_t)FQ_SCHOOL_WAVE,
                                            (uint32_t)(pd.items_off + k * pd.item_size), m, FQ_NONE,
                                            {FQ_NONE, FQ_NONE}, {0, 0}});
                }
                continue;
            }
            uint32_t qk;
            if (kind == 0) qk = FQ_OUTER | (pd.L.inv ? 1u : 0u) | (pd.L.top ? 2u : 0u) | (pd.L.r3 ? 4u : 0u);
            else if (kind == 1) qk = !pd.L.small ? FQ_INNER : (pd.L.r3 ? FQ_INNER_SMALL_R3 : FQ_INNER_SMALL);
            else qk = FQ_CRT;
            qlds = std::max(qlds, kind == 2 ? ff_queue_lds_crt() : pd.L.lds);
            std::vector<uint32_t> added(ncnt, 0);
            for (size_t k = 0; k < pd.who.size(); k++) {
                const int c = pd.who[k];
                const uint32_t end = k + 1 < pd.start.size() ? pd.start[k + 1] : pd.L.blocks;
                FFQTask t{qk, (uint32_t)(pd.items_off + k * pd.item_size), 0, kind == 2 ? FQ_NONE : (uint32_t)c,
                          {FQ_NONE, FQ_NONE}, {0, 0}};
                // the product's tasks in earlier launches: its previous passes
                if (count[c]) t.dep[0] = (uint32_t)c, t.need[0] = count[c];
                // inner blocks also read a b transform made by another product
                if (kind == 1 && c < (int)n) {
                    if (b_owner[c] >= 0) t.dep[1] = (uint32_t)b_owner[c], t.need[1] = fwd_total[b_owner[c]];
                    else if (pre_of[c] >= 0) t.dep[1] = (uint32_t)(n + pre_of[c]), t.need[1] = 3;
                }
                for (uint32_t b = pd.start[k]; b < end; b++) {
                    t.sub = b - pd.start[k];
                    tasks.push_back(t);
                }
                added[c] += end - pd.start[k];
            }
            for (size_t c = 0; c < ncnt; c++) count[c] += added[c];
        }
        for (size_t i = 0; i < n; i++) {
            if (!B->zero[i]) continue;
            const uint64_t *out = B->items[i].out;
            tasks.push_back(FFQTask{FQ_FILL, (uint32_t)put(&out, sizeof out), 0, FQ_NONE, {FQ_NONE, FQ_NONE}, {0, 0}});
        }
        tasks_off = put(tasks.data(), tasks.size() * sizeof(FFQTask));
        B->ntasks = (uint32_t)tasks.size();
        B->nctr = (uint32_t)ncnt;
        B->qlds = qlds;
        B->qbytes = qbytes;
        B->ctr = (uint32_t *)ff_dalloc((FQ_HEADER + ncnt) * sizeof(uint32_t));
        FF_CHECK(hipMemsetAsync(B->ctr, 0, (FQ_HEADER + ncnt) * sizeof(uint32_t), g_st));
    }
    if (img.size() > 0xffffffffull) {
        fprintf(stderr, "fastfib_amd: batch descriptors past 4 GiB\n");
        exit(1);
    }
    B->desc = (char *)ff_dalloc(img.size() + 256);
    ff_copy(B->desc, img.data(), img.size(), hipMemcpyHostToDevice);
    if (B->flags & FF_BATCH_QUEUE) {
        B->qtasks = (const FFQTask *)(B->desc + tasks_off);
        return;
    }
    for (auto &pd : pend) {
        BLaunch L = pd.L;
        L.items = B->desc + pd.items_off;
        L.start = (const uint32_t *)(B->desc + pd.start_off);
        B->launches.push_back(L);
    }
    for (size_t i = 0; i < n; i++) {
        if (!B->zero[i]) continue;
        BLaunch L{};
        L.kind = 3;
        L.fill = B->items[i].out;
        B->launches.push_back(L);
    }
}

static void enqueue_grouped(ff_batch *B) {
    if (B->qtasks) {
        FF_CHECK(ff_launch_queue(B->qtasks, B->ntasks, B->desc, B->ctr, B->nctr, B->qlds, B->qbytes, g_st));
        return;
    }
    for (const BLaunch &L : B->launches) {
        switch (L.kind) {
        case 0:
            FF_CHECK(ff_launch_outer_list(L.inv, L.top, L.r3, L.sh, (const FFOuterItem *)L.items, L.start, L.n, L.blocks,
                                          L.lds, L.bytes, g_st));
            break;
        case 1:
            FF_CHECK(ff_launch_inner_list(L.small, L.r3, L.sh != 0, (const FFInnerItem *)L.items, L.start, L.n, L.blocks, L.lds,
                                          L.bytes, g_st));
            break;
        case 2:
            FF_CHECK(ff_launch_crt_list((const FFCrtJob *)L.items, L.start, L.n, L.blocks, L.conv_total, L.out_total,
                                        g_st));
            break;
        case 4:
            FF_CHECK(ff_launch_school_list((const FFSchoolItem *)L.items, L.n, L.bytes, g_st));
            break;
        case 5:
            FF_CHECK(ff_launch_school_wave_list((const FFSchoolItem *)L.items, L.n, L.bytes, g_st));
            break;
        default:
            FF_CHECK(ff_launch_fill(L.fill, 1, 0, g_st));
            break;
        }
    }
}

extern "C" int ff_batch_create(ff_batch **out, const ff_prod *items, size_t count, int flags) {
    return ff_batch_create_sched(out, items, count, flags, nullptr);
}

int ff_batch_create_sched(ff_batch **out, const ff_prod *items, size_t count, int flags, const FFSched *force) {
    if (!out || (!items && count)) return FF_ERR_ARG;
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    ff_ensure_init();
    ff_batch *B = new ff_batch;
    B->flags = flags;
    B->items.assign(items, items + count);
    for (const ff_prod &p : B->items) {
        if (!p.a || !p.out || p.na == 0 || (p.b && p.b != p.a && p.nb == 0)) {
            delete B;
            ff_set_err("batch item without operands or output");
            return FF_ERR_ARG;
        }
    }
    read_bits(B->items);
    B->shape.resize(count);
    B->sched.resize(count);
    B->zero.assign(count, 0);
    B->school.assign(count, 0);
    std::vector<ProdShape> s0;
    std::vector<uint64_t> ba, bb;
    std::vector<uint8_t> sq;
    std::vector<size_t> who;
    for (size_t i = 0; i < count; i++) {
        ff_prod &p = B->items[i];
        const bool square = !p.b || p.b == p.a;
        const uint64_t bits_b = square ? p.bits_a : p.bits_b;
        if (p.bits_a == 0 || bits_b == 0) {
            B->zero[i] = 1;
            continue;
        }
        if (ff_school_ok(p.na, square ? p.na : p.nb)) B->school[i] = 1;
        const ProdShape s = ff_product_shape(p.bits_a, bits_b);
        if (s.split) {
            delete B;
            ff_set_err("batch item %zu needs the split multiply", i);
            return FF_ERR_UNSUPPORTED;
        }
        s0.push_back(s), ba.push_back(p.bits_a), bb.push_back(bits_b), sq.push_back(square), who.push_back(i);
    }
    std::vector<FFSched> pick(who.size());
    if (force) {
        for (size_t j = 0; j < who.size(); j++) pick[j] = force[who[j]];
    } else {
        ff_pick_scheds(s0.data(), ba.data(), bb.data(), sq.data(), who.size(), count, flags, pick.data());
    }
    for (size_t j = 0; j < who.size(); j++) {
        B->sched[who[j]] = pick[j];
        B->shape[who[j]] = ff_shape_with(s0[j], pick[j]);
    }
    B->slots = (unsigned long long *)ff_dalloc((count ? count : 1) * 8);
    if (!(flags & (FF_BATCH_GROUPED | FF_BATCH_GRAPH | FF_BATCH_QUEUE))) {
        for (size_t i = 0; i < count; i++) {
            if (B->zero[i]) continue;
            const ff_prod &p = B->items[i];
            const bool sq = !p.b || p.b == p.a;
            B->plain.push_back(MulItem{p.a, p.na, p.bits_a, sq ? nullptr : p.b, sq ? p.na : p.nb,
                                       sq ? p.bits_a : p.bits_b, p.out, B->shape[i].conv, B->slots + i});
        }
    } else {
        build_lists(B);
        if (flags & FF_BATCH_GRAPH) {
            FF_CHECK(hipStreamSynchronize(g_st));
            ff_prof_pause(1);
            FF_CHECK(hipStreamBeginCapture(g_st, hipStreamCaptureModeRelaxed));
            enqueue_grouped(B);
            FF_CHECK(hipStreamEndCapture(g_st, &B->graph));
            ff_prof_pause(0);
            FF_CHECK(hipGraphInstantiate(&B->exec, B->graph, nullptr, nullptr, 0));
        }
    }
    *out = B;
    return FF_OK;
}

extern "C" int ff_batch_run(ff_batch *B) {
    if (!B) return FF_ERR_ARG;
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    if (B->exec) {
        FF_CHECK(hipGraphLaunch(B->exec, g_st));
    } else if (!B->launches.empty() || B->qtasks || (B->flags & (FF_BATCH_GROUPED | FF_BATCH_QUEUE))) {
        enqueue_grouped(B);
    } else {
        size_t k = 0;
        for (size_t i = 0; i < B->items.size(); i++) {
            if (B->zero[i]) {
                FF_CHECK(ff_launch_fill(B->items[i].out, 1, 0, g_st));
                continue;
            }
            const MulItem &m = B->plain[k++];
            if (B->school[i]) {
                FFSchoolSet ss = {};
                ss.j[0] = FFSchoolItem{m.a, m.b ? m.b : m.a, (uint32_t)m.na, (uint32_t)(m.b ? m.nb : m.na), m.out,
                                       m.lenkey};
                FF_CHECK(ff_launch_school_set(&ss, 1, 8.0 * (2 * m.na + 2 * (m.b ? m.nb : m.na) + 1), g_st));
                continue;
            }
            ff_mul_group(&m, 1, B->shape[i], B->sched[i]);
        }
    }
    return FF_OK;
}

extern "C" int ff_batch_finish(ff_batch *B, ff_prod *items) {
    if (!B) return FF_ERR_ARG;
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    const size_t n = B->items.size();
    std::vector<uint64_t> keys(n ? n : 1);
    uint32_t fault = 0;
    if (n) FF_CHECK(hipMemcpyAsync(keys.data(), B->slots, n * 8, hipMemcpyDeviceToHost, g_st));
    FF_CHECK(hipMemcpyAsync(&fault, ff_lb_fault(), 4, hipMemcpyDeviceToHost, g_st));
    FF_CHECK(hipStreamSynchronize(g_st));
    if (fault) ff_lb_timeout_fail();
    if (items) {
        for (size_t i = 0; i < n; i++) items[i].nout = B->zero[i] || keys[i] == 0 ? 1 : (size_t)(keys[i] >> 8);
    }
    return FF_OK;
}

extern "C" void ff_batch_destroy(ff_batch *B) {
    if (!B) return;
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    (void)hipStreamSynchronize(g_st);
    if (B->exec) (void)hipGraphExecDestroy(B->exec);
    if (B->graph) (void)hipGraphDestroy(B->graph);
    for (void *w : B->work) ff_dfree(w);
    ff_dfree(B->desc);
    ff_dfree(B->slots);
    if (B->ctr) ff_dfree(B->ctr);
    delete B;
}
