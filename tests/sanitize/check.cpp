// Host code under AddressSanitizer and UndefinedBehaviorSanitizer: the
// oracle (store, batched updates, compaction, every walk app, export, float
// weights, dense groups) and the library's host-side generators (R-MAT arcs, update
// streams with deletes, expiry and compaction renumbering).
// Built and run by tests/test_sanitize_cpu.py; any report aborts the run.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../include/bingo.h"
#include "../../oracle/bingo_oracle.h"

#define CHECK(x)                                                          \
    do {                                                                  \
        if (!(x)) {                                                       \
            std::fprintf(stderr, "check failed: %s (line %d)\n", #x, __LINE__); \
            std::exit(1);                                                 \
        }                                                                 \
    } while (0)

static void export_all(orc_graph *g)
{
    orc_dump d{};
    CHECK(orc_export(g, &d) == ORC_OK);
    const int32_t n = orc_num_nodes(g);
    std::vector<uint32_t> nslots(n + 1), bsize((size_t)(n + 1) * 32), slot_v(d.total_slots + 1),
        slot_w(d.total_slots + 1), backpos(d.total_members + 1);
    std::vector<uint64_t> mass(n + 1), T((size_t)(n + 1) * 32), members(d.total_members + 1);
    std::vector<uint8_t> alias((size_t)(n + 1) * 32);
    std::vector<uint32_t> dmask(n + 1);
    d.dmask = dmask.data();
    d.nslots = nslots.data();
    d.mass = mass.data();
    d.bsize = bsize.data();
    d.T = T.data();
    d.alias = alias.data();
    d.slot_v = slot_v.data();
    d.slot_w = slot_w.data();
    d.members = members.data();
    d.backpos = backpos.data();
    CHECK(orc_export(g, &d) == ORC_OK);
}

static void walks(orc_graph *g, int fw)
{
    const int32_t n = orc_num_nodes(g);
    for (int app = 0; app < 4; ++app) {
        orc_walk_params p{};  // float stores take node2vec's plain envelope (no folding)
        orc_walk_thresholds_ex(0.5, 2.0, 0.15, fw ? 0 : 1, &p);
        p.seed = 11 + app;
        p.max_len = 24;
        p.max_attempts = 64;
        std::vector<int32_t> paths((size_t)n * p.max_len);
        uint64_t steps = 0;
        CHECK(orc_walk(g, app, &p, nullptr, nullptr, n, paths.data(), &steps, 2) == ORC_OK);
    }
}

int main()
{
    const int scale = 9;
    const int32_t n = 1 << scale;
    uint32_t thr[3];
    orc_rmat_thresholds(0.57, 0.19, 0.19, thr);
    // integer weights, float weights, integer weights with dense groups
    for (int mode = 0; mode < 3; ++mode) {
        const int fw = mode == 1;
        const int64_t m = 6000;
        std::vector<orc_edge> base((size_t)m * 2);
        orc_gen_rmat(scale, 0, m, thr, ORC_W_LOGUNI32, 5, 1, base.data(), 2);
        if (fw)
            for (auto &e : base) {
                const float x = (float)e.w / 256.0f;
                std::memcpy(&e.w, &x, 4);
            }
        orc_graph *g = orc_create_rate(n, 1, fw, mode == 2 ? 26214u : 0u);
        CHECK(orc_insert(g, base.data(), (int64_t)base.size()) == ORC_OK);
        // the library's generator drives the updates (expiry, then compaction)
        bingo_updgen *ug = nullptr;
        std::vector<bingo_edge> lb(base.size());
        for (size_t i = 0; i < base.size(); ++i) lb[i] = bingo_edge{base[i].u, base[i].v, base[i].w};
        CHECK(bingo_updgen_create(n, scale, lb.data(), (int64_t)lb.size(), m, 0.57, 0.19, 0.19, BINGO_W_LOGUNI32, 5,
                                  &ug) == BINGO_OK);
        for (int r = 0; r < 6; ++r) {
            std::vector<bingo_edge> ins(900);
            std::vector<bingo_del> del(900);
            CHECK(bingo_updgen_inserts(ug, 900, ins.data()) == BINGO_OK);
            CHECK(bingo_updgen_expire(ug, 900, del.data()) == BINGO_OK);
            std::vector<orc_edge> oi(900);
            for (int i = 0; i < 900; ++i) {
                oi[i] = orc_edge{ins[i].u, ins[i].v, ins[i].w};
                if (fw) {
                    const float x = (float)ins[i].w / 256.0f;
                    std::memcpy(&oi[i].w, &x, 4);
                }
            }
            CHECK(orc_insert(g, oi.data(), 900) == ORC_OK);
            CHECK(orc_delete(g, reinterpret_cast<const orc_del *>(del.data()), 900) == ORC_OK);
            const double frac = r % 2 ? 0.05 : 0.0;
            CHECK(bingo_updgen_compact(ug, frac) == BINGO_OK);
            uint64_t freed = 0;
            CHECK(orc_compact(g, frac, &freed) >= 0);
            export_all(g);
        }
        walks(g, fw);
        bingo_updgen_destroy(ug);
        orc_destroy(g);
    }
    // random deletes and the general renumbering path
    {
        const int64_t m = 4000;
        std::vector<bingo_edge> base((size_t)m);
        CHECK(bingo_gen_rmat(scale, 0, m, 0.57, 0.19, 0.19, BINGO_W_UNIFORM16, 8, 0, base.data()) == BINGO_OK);
        bingo_updgen *ug = nullptr;
        CHECK(bingo_updgen_create(n, scale, base.data(), m, m, 0.57, 0.19, 0.19, BINGO_W_UNIFORM16, 8, &ug) == BINGO_OK);
        std::vector<bingo_del> del(700);
        std::vector<bingo_edge> ins(500);
        for (int r = 0; r < 5; ++r) {
            CHECK(bingo_updgen_inserts(ug, 500, ins.data()) == BINGO_OK);
            CHECK(bingo_updgen_deletes(ug, 700, del.data()) == BINGO_OK);
            CHECK(bingo_updgen_compact(ug, 0.1) == BINGO_OK);
        }
        bingo_updgen_destroy(ug);
    }
    std::printf("sanitized host checks ok\n");
    return 0;
}
