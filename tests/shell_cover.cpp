// Host-side check of the fused step's CPML shell decomposition (shell_boxes.hpp, the
// code the kernels use): for grids, CPML widths and line phases, the shell boxes decoded
// thread by thread (shell_quad) cover every quad outside the interior box exactly once,
// never a quad inside it, and every decoded (i, j, k0) lies inside the grid; the same
// for the two-step box with its low and high fringes added.  Exit status 0 when all hold.
#include "../fdtd_accelerate_amd/csrc/shell_boxes.hpp"

#include <cstdio>
#include <vector>

using namespace mifdtd;

static int check(int nx, int ny, int nz, int nzp, int npml, int ph, bool pairs)
{
    const int Q = nzp / 4;
    FGeom g{};
    g.nx = nx; g.ny = ny; g.nzp = nzp; g.nz = nz;
    const int d = pairs ? 2 : 1;  // the box keeps d cells to the CPML slabs
    g.xb = npml + d; g.xe = nx - d - npml;
    g.yb = npml + d; g.ye = ny - d - npml;
    g.kb = line_up(npml + d, ph); g.ke = line_down(nz - d - npml, ph);
    if (g.xb >= g.xe || g.yb >= g.ye || g.kb >= g.ke) return 0;  // no interior box: nothing to check
    if (pairs && (g.xe - g.xb < 4 || g.ye - g.yb < 4)) return 0;
    ShellBoxes sb{};
    unsigned long long total = 0;
    add_shell(sb, total, g, Q);
    if (pairs) {
        add_low_fringe(sb, total, g);
        add_high_fringe(sb, total, g);
    }
    if (sb.n > kMaxShellBoxes) { std::printf("box overflow %d\n", sb.n); return 1; }
    sb.start[sb.n] = (unsigned)total;
    std::vector<int> hits((size_t)nx * ny * Q, 0);
    for (unsigned long long f = 0; f < total; ++f) {
        int i, j, k0;
        shell_quad(sb, (unsigned)f, i, j, k0);
        if (i < 0 || i >= nx || j < 0 || j >= ny || k0 < 0 || k0 >= nzp || k0 % 4) {
            std::printf("out of bounds: thread %llu -> (%d, %d, %d) in (%d, %d, %d)\n", f, i, j, k0, nx, ny, nzp);
            return 1;
        }
        ++hits[((size_t)i * ny + j) * Q + k0 / 4];
    }
    const int qb = g.kb / 4, qe = g.ke / 4;
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int q = 0; q < Q; ++q) {
                const bool in_box = i >= g.xb && i < g.xe && j >= g.yb && j < g.ye && q >= qb && q < qe;
                // the fringes: the box's first plane, row and quad, and its last plane, row and quad
                const bool fringe = pairs && in_box &&
                                    (i == g.xb || j == g.yb || q == qb || i == g.xe - 1 || j == g.ye - 1 || q == qe - 1);
                const int want = (!in_box || fringe) ? 1 : 0;
                const int got = hits[((size_t)i * ny + j) * Q + q];
                if (got != want) {
                    std::printf("quad (%d, %d, %d) of (%d, %d, %d), npml %d, phase %d%s: covered %d times, expected %d\n",
                                i, j, 4 * q, nx, ny, nzp, npml, ph, pairs ? ", pairs" : "", got, want);
                    return 1;
                }
            }
    return 0;
}

int main()
{
    const int grids[][3] = {{24, 20, 32}, {30, 26, 48}, {29, 31, 64}, {40, 36, 160}, {17, 45, 544},
                            {26, 70, 300}, {64, 64, 64}, {33, 40, 416}, {12, 13, 1008}, {50, 48, 128}};
    int checked = 0;
    for (const auto &gr : grids)
        for (int npml = 0; npml <= 12; npml += 3)
            for (int ph = 0; ph < 32; ph += 4)
                for (int pairs = 0; pairs < 2; ++pairs) {
                    const int nz = gr[2] - (gr[2] % 16 == 0 ? 0 : gr[2] % 16);
                    if (pairs && npml == 0) continue;
                    if (int rc = check(gr[0], gr[1], nz, gr[2], npml, gr[2] % 32 ? 0 : ph, pairs != 0)) return rc;
                    ++checked;
                }
    std::printf("shell cover ok: %d cases\n", checked);
    return 0;
}
