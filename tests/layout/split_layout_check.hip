// CPU check of the row-block layout (engine/devmat.hpp, segment_blocks) that
// the pdlp kernels rely on when they split rows of more than kLongRow
// nonzeros over workgroups: every nonzero is covered once and in order,
// pieces start on a group boundary of their row, hold at most
// max(tile, kGroupLen) nonzeros, end in a partial group only at the row's
// end, and write only their row's group slots; split rows have exactly the
// pieces their table entry counts; nothing is split when splitting is off.
// Host code only (no device call). Exit status 0 = all checks hold.
#include <cstdio>
#include <random>
#include <vector>

#include "../../nirnaya_amd/csrc/engine/devmat.hpp"

using namespace nir::gpu;

int main() {
    std::mt19937 rng(3);
    long long bad = 0, rows_split = 0;
    for (int trial = 0; trial < 300; ++trial) {
        const int nrows = 1 + static_cast<int>(rng() % 50);
        std::vector<int> rp{0};
        for (int r = 0; r < nrows; ++r) {
            int len = rng() % 4 == 0 ? 4090 + static_cast<int>(rng() % 20000) : static_cast<int>(rng() % 300);
            if (trial % 7 == 0 && r == 0) len = 600000;
            rp.push_back(rp.back() + len);
        }
        for (int tile : {1024, 2048, 4096}) {
            for (int split : {0, 1}) {
                TileShape sh;
                sh.tile = tile;
                sh.split = split;
                std::vector<int4> desc, lrow;
                int groups = 0;
                segment_blocks(rp.data(), nrows, sh, split != 0, desc, lrow, groups);
                if (!split && !lrow.empty()) ++bad;
                int pos = 0, gs = 0;
                std::vector<int> seen(lrow.size(), 0);
                for (const int4& L : lrow) {
                    if (L.x != gs) ++bad;
                    gs += L.y;
                }
                if (gs != groups) ++bad;
                for (const int4& d : desc) {
                    if (d.z != pos) ++bad;   // nonzeros covered once, in order
                    pos = d.w;
                    if (d.y >= 0) {
                        if (d.w - d.z > tile && d.y != d.x + 1) ++bad;   // only single rows exceed a tile
                        if (split && d.y == d.x + 1 && d.w - d.z > kLongRow) ++bad;   // long rows are split
                        continue;
                    }
                    const int id = -d.y - 1;
                    if (id < 0 || id >= static_cast<int>(lrow.size())) {
                        ++bad;
                        continue;
                    }
                    const int4 L = lrow[id];
                    ++seen[id];
                    const int r0 = rp[L.w], r1 = rp[L.w + 1];
                    if (d.x != L.w || r1 - r0 <= kLongRow) ++bad;
                    if ((d.z - r0) % kGroupLen) ++bad;
                    if (d.w - d.z > std::max(tile, kGroupLen)) ++bad;
                    if ((d.w - d.z) % kGroupLen && d.w != r1) ++bad;
                    const int g0 = L.x + (d.z - r0) / kGroupLen;
                    const int g1 = g0 + (d.w - d.z + kGroupLen - 1) / kGroupLen;
                    if (g0 < L.x || g1 > L.x + L.y) ++bad;
                    if ((r1 - r0 + kGroupLen - 1) / kGroupLen != L.y) ++bad;
                }
                if (pos != rp.back()) ++bad;
                for (size_t i = 0; i < lrow.size(); ++i)
                    if (seen[i] != lrow[i].z) ++bad;
                rows_split += static_cast<long long>(lrow.size());
            }
        }
    }
    std::printf("split layout: %lld split rows checked, %lld violations\n", rows_split, bad);
    return bad == 0 ? 0 : 1;
}
