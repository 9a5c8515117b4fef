// Host-side check of the register march's dispatch order (edge_first_tile): for many grid shapes,
// group mixes and tile shapes, every block maps to a distinct tile of the launch, the group it
// reports owns that tile, and the first nedge blocks take exactly the tiles that touch an edge.
// Built host-only against the library's own header by tests/test_tile_order.py.
#include <cstdio>
#include <vector>
#include "swe2d_desc.hpp"

using namespace his;

static bool touches_edge(int nx, int ny, int tx, int ty, int s, int r) {
    const int x0 = s * tx, y0 = r * ty;
    return !(x0 >= 3 && x0 + tx + 3 <= nx && y0 >= 3 && y0 + ty + 3 <= ny);
}

static int check(const std::vector<int>& nxs, const std::vector<int>& nys, const std::vector<int>& Bs, int tx, int ty) {
    Launch2D<float> L{};
    L.ngroups = (int)nxs.size();
    for (int gi = 0; gi < L.ngroups; ++gi) { L.g[gi].nx = nxs[gi]; L.g[gi].ny = nys[gi]; L.g[gi].B = Bs[gi]; }
    tiles_for(L, tx, ty);
    std::vector<int> hit(L.ntiles, 0);
    for (int p = 0; p < L.ntiles; ++p) {
        int gi = -1;
        const int t = edge_first_tile<float>(L, p, gi);
        if (t < 0 || t >= L.ntiles) { std::printf("FAIL block %d -> tile %d of %d\n", p, t, L.ntiles); return 1; }
        if (gi < 0 || gi >= L.ngroups || t < L.g[gi].tile0 ||
            t >= L.g[gi].tile0 + L.g[gi].B * L.g[gi].tiles_x * L.g[gi].tiles_y) {
            std::printf("FAIL block %d -> tile %d outside group %d\n", p, t, gi); return 1;
        }
        const Group2D<float>& G = L.g[gi];
        const int rem = (t - G.tile0) % (G.tiles_x * G.tiles_y);
        const bool e = touches_edge(G.nx, G.ny, tx, ty, rem % G.tiles_x, rem / G.tiles_x);
        if (e != (p < L.nedge)) { std::printf("FAIL block %d (nedge %d) -> tile %d edge=%d\n", p, L.nedge, t, (int)e); return 1; }
        hit[t] += 1;
    }
    for (int t = 0; t < L.ntiles; ++t)
        if (hit[t] != 1) { std::printf("FAIL tile %d hit %d times\n", t, hit[t]); return 1; }
    return 0;
}

int main() {
    const int shapes[][2] = {{58, 8}, {58, 10}, {58, 12}, {58, 13}, {58, 16}, {58, 32}, {58, 4},
                             {32, 8}, {64, 4}, {16, 8}, {32, 4}, {64, 8}, {64, 16}};
    const int sizes[][2] = {{400, 325}, {1, 1}, {5, 7}, {58, 13}, {61, 19}, {64, 64}, {65, 20}, {120, 9},
                            {200, 200}, {117, 333}, {300, 17}, {3, 500}, {500, 3}, {131, 97}};
    int n = 0;
    for (const auto& sh : shapes) {
        for (const auto& a : sizes) {
            for (int B : {1, 3, 32}) {
                if (check({a[0]}, {a[1]}, {B}, sh[0], sh[1])) return 1;
                ++n;
                for (const auto& b : sizes) {   // two groups of different grids
                    if (check({a[0], b[0]}, {a[1], b[1]}, {B, 2}, sh[0], sh[1])) return 1;
                    ++n;
                }
            }
        }
    }
    // three groups
    if (check({400, 120, 65}, {325, 9, 20}, {4, 1, 7}, 58, 13)) return 1;
    std::printf("ok %d launches\n", n + 1);
    return 0;
}
