// Host check of the block-to-item mapping of every striped launch in
// libmpas_sw_amd (msw_launch.hpp, compiled here for the CPU): for each
// launch family, grid size, stripe option, block size and batch it walks
// every hardware thread of the grid and checks that
//   * the live threads cover the launch's items exactly once and no live
//     thread's item lies past the last one (so every row a kernel touches,
//     rowp(p, item >> lb, B)[item & (B-1)], is inside its [count][B] buffer);
//   * with B >= 64 (one wavefront per entity), the 64 lanes of a wavefront
//     are all live or all dead, share one entity (the wave-uniform scalar
//     indexing) and lane L holds member L mod 64 (the pair form's swap);
//   * in the 16-lanes-per-edge coefficient launch, each 16-lane group is all
//     live or all dead and lane L holds slot L mod 16 (its shuffles).
// Usage: launch_map_check [quick]; prints one line per mesh and exits 1 on
// the first failure, naming the launch.
#define MSW_HD
#include "../mpas-sw-tl-ad_amd/csrc/msw_launch.hpp"

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

using namespace msw::launch;

namespace {

struct Mesh {
    const char* name;
    int nC, nE, nV;
};

long g_threads = 0;

[[noreturn]] void die(const char* what, const char* mesh, int lb, int sl, int block, long detail)
{
    std::printf("FAIL %s mesh=%s lb=%d sl=%d block=%d detail=%ld\n", what, mesh, lb, sl, block, detail);
    std::exit(1);
}

// a flat launch of `total` items (entity launches: total = count << lb)
void check_flat(const char* what, const Mesh& m, unsigned total, int lb, int sl, int block, int group)
{
    const int grid = grid_items(static_cast<long>(total), sl, block);
    std::vector<unsigned char> seen(total, 0);
    for (unsigned bid = 0; bid < static_cast<unsigned>(grid); ++bid) {
        for (int w = 0; w < block; w += group) {
            // one group: a 64-lane wavefront (B >= 64) or a 16-lane slot group
            int live = 0;
            unsigned ent = 0;
            for (int l = 0; l < group; ++l) {
                const unsigned tid = static_cast<unsigned>(w + l);
                const Item it = flat_item(bid, tid, static_cast<unsigned>(grid), block, sl, total);
                ++g_threads;
                if (!it.live) continue;
                if (it.t >= total) die(what, m.name, lb, sl, block, it.t);
                if (seen[it.t]++) die(what, m.name, lb, sl, block, -static_cast<long>(it.t));
                if (group > 1) {
                    if ((it.t & static_cast<unsigned>(group - 1)) != (tid & static_cast<unsigned>(group - 1)))
                        die("lane/member", m.name, lb, sl, block, it.t);
                    const unsigned e = group == 64 ? it.t >> lb : it.t >> 4;
                    if (live && e != ent) die("wave-uniform entity", m.name, lb, sl, block, it.t);
                    ent = e;
                }
                ++live;
            }
            if (group > 1 && live != 0 && live != group) die("partial group", m.name, lb, sl, block, bid);
        }
    }
    for (unsigned i = 0; i < total; ++i)
        if (!seen[i]) die(what, m.name, lb, sl, block, static_cast<long>(i) + 1000000000L);
}

void check_edge_cell(const Mesh& m, int lb, int sl, int block)
{
    const int grid = grid_edge_cell(m.nE, m.nC, lb, sl, block);
    const unsigned tE = static_cast<unsigned>(m.nE) << lb, tC = static_cast<unsigned>(m.nC) << lb;
    std::vector<unsigned char> se(tE, 0), sc(tC, 0);
    const int group = lb >= 6 ? 64 : 1;
    for (unsigned bid = 0; bid < static_cast<unsigned>(grid); ++bid)
        for (int w = 0; w < block; w += group) {
            int live = 0;
            unsigned ent = 0;
            bool kind = false;
            for (int l = 0; l < group; ++l) {
                const unsigned tid = static_cast<unsigned>(w + l);
                const Item it = edge_cell_item(bid, tid, static_cast<unsigned>(grid), block, sl, m.nE, m.nC, lb);
                ++g_threads;
                if (!it.live) continue;
                std::vector<unsigned char>& s = it.is_edge ? se : sc;
                if (it.t >= (it.is_edge ? tE : tC)) die("edge/cell range", m.name, lb, sl, block, it.t);
                if (s[it.t]++) die("edge/cell duplicate", m.name, lb, sl, block, it.t);
                if (group > 1) {
                    if ((it.t & 63u) != (tid & 63u)) die("edge/cell lane/member", m.name, lb, sl, block, it.t);
                    if (live && ((it.t >> lb) != ent || it.is_edge != kind))
                        die("edge/cell wave-uniform entity", m.name, lb, sl, block, it.t);
                    ent = it.t >> lb;
                    kind = it.is_edge;
                }
                ++live;
            }
            if (group > 1 && live != 0 && live != group) die("edge/cell partial wave", m.name, lb, sl, block, bid);
        }
    for (unsigned i = 0; i < tE; ++i)
        if (!se[i]) die("edge item missed", m.name, lb, sl, block, i);
    for (unsigned i = 0; i < tC; ++i)
        if (!sc[i]) die("cell item missed", m.name, lb, sl, block, i);
}

void check_mesh(const Mesh& m, const std::vector<int>& lbs, const std::vector<int>& sls,
                const std::vector<int>& blocks, bool coef)
{
    for (int lb : lbs)
        for (int sl : sls)
            for (int block : blocks) {
                const int group = lb >= 6 ? 64 : 1;
                check_flat("cell items", m, static_cast<unsigned>(m.nC) << lb, lb, sl, block, group);
                check_flat("edge items", m, static_cast<unsigned>(m.nE) << lb, lb, sl, block, group);
                check_flat("vertex items", m, static_cast<unsigned>(m.nV) << lb, lb, sl, block, group);
                check_edge_cell(m, lb, sl, block);
                if (coef && lb == lbs.front()) {
                    // B = 1 coefficient launches (k_tl_coef, k_adj_coef_*)
                    check_flat("tl_coef items", m, 16u * m.nE + 8u * m.nC, 0, sl, block, 1);
                    check_flat("adj_coef_edge items", m, 16u * m.nE, 0, sl, block, 16);
                    check_flat("adj_coef_vertex items", m, 8u * m.nV, 0, sl, block, 1);
                    check_flat("adj_coef_cell items", m, 8u * m.nC, 0, sl, block, 1);
                }
            }
    std::printf("ok %s (C=%d E=%d V=%d)\n", m.name, m.nC, m.nE, m.nV);
    std::fflush(stdout);
}

}  // namespace

int main(int argc, char** argv)
{
    const bool quick = argc > 1 && std::strcmp(argv[1], "quick") == 0;
    const std::vector<int> all_sl = {0, 1, 2, 3, 4, 5, 6, 7}, all_blk = {64, 128, 256, 512};
    // icosahedral levels 0, 3 (the GPU tests' mesh) and 5, every batch size
    check_mesh({"L0", 12, 30, 20}, {0, 1, 2, 5, 6, 7, 8}, all_sl, all_blk, true);
    check_mesh({"L3", 642, 1920, 1280}, {0, 2, 6, 7, 8}, all_sl, all_blk, true);
    check_mesh({"L5", 10242, 30720, 20480}, {0, 6}, all_sl, all_blk, true);
    if (!quick) {
        // level 7 with 64 members: every stripe at 256 threads, every block
        // size at the default stripe and at one range per XCD
        const Mesh l7 = {"L7", 163842, 491520, 327680};
        check_mesh(l7, {6}, all_sl, {256}, true);
        check_mesh(l7, {6}, {2, 7}, {64, 128, 512}, true);
        // the uneven batch 64 x L5 + 16 x L6 + 4 x L7 + 1 x L9, one member
        check_mesh({"union", 4587690, 13762560, 9175040}, {0}, {2, 7}, {256}, false);
    }
    std::printf("launch map ok: %ld threads checked\n", g_threads);
    return 0;
}
