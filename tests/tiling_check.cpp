// Host check of the column tiling of the z-marching kernels
// (dumses-hybrid_amd/csrc/dmx_tiling.h): for many domain sizes and the
// tile limits the kernels use, every column and plane lies in exactly one
// tile, and every tile respects its width, height, column and plane
// limits.  Prints the tilings of a 256^3 domain and "bad 0" on success.
#include <cstdio>
#include <initializer_list>
#include <vector>

#include "dmx_tiling.h"

using namespace dmx;

int main()
{
    long bad = 0, cases = 0;
    // (widest tile, highest tile, columns per block, nominal planes):
    // the fused kernel in float32 and float64, a square variant, kt = 1
    const int params[][4] = {{33, 8, 256, 16}, {33, 8, 256, 8}, {33, 8, 256, 32}, {17, 8, 128, 16},
                             {17, 16, 256, 8}, {33, 8, 256, 1}};
    for (const auto& pr : params) {
        const int fxm = pr[0], fym = pr[1], fc = pr[2], kt = pr[3];
        for (int nx = 1; nx <= 300; nx += (nx < 70 ? 1 : 7))
            for (int ny = 1; ny <= 140; ny += (ny < 40 ? 1 : 13))
                for (int nz : {1, 2, 5, 17, 33, 100, 258}) {
                    if (nz >= 100 && (nx % 5 || ny % 4)) continue;   // keep the run to seconds
                    const Tiling t = make_tiling(nx, ny, nz, fxm, fym, fc, kt);
                    std::vector<int> cov((size_t)nx * ny * nz, 0);
                    for (int zb = 0; zb < t.pz.nt; zb++)
                        for (int l = 0; l < t.tz; l++) {
                            int x0, fx, y0, fy, z0, fz;
                            tiling_tile(t, l, zb, x0, fx, y0, fy, z0, fz);
                            if (fx < 1 || fx > fxm || fy < 1 || fy > fym || fx * fy > fc || fz < 1 ||
                                (fz < kt && fz != nz) || x0 < 0 || y0 < 0 || z0 < 0 || x0 + fx > nx ||
                                y0 + fy > ny || z0 + fz > nz) {
                                if (bad++ < 5)
                                    printf("bad tile: n %d %d %d limits %d %d %d kt %d: x %d+%d y %d+%d z %d+%d\n", nx,
                                           ny, nz, fxm, fym, fc, kt, x0, fx, y0, fy, z0, fz);
                                continue;
                            }
                            for (int z = z0; z < z0 + fz; z++)
                                for (int y = y0; y < y0 + fy; y++)
                                    for (int x = x0; x < x0 + fx; x++) cov[((size_t)z * ny + y) * nx + x]++;
                        }
                    for (size_t c = 0; c < cov.size(); c++)
                        if (cov[c] != 1) {
                            if (bad++ < 5) printf("column covered %d times: n %d %d %d\n", cov[c], nx, ny, nz);
                            break;
                        }
                    cases++;
                }
    }
    const Tiling f = make_tiling(257, 257, 257, 33, 8, 256, 16);
    printf("faces of 256^3, float32 fused kernel: %d x strips (%d of %d, %d of %d), %d tiles per block of "
           "planes, %d blocks of planes\n",
           f.px.nt, f.px.nt - f.px.nwide, f.px.w, f.px.nwide, f.px.w + 1, f.tz, f.pz.nt);
    printf("cases %ld bad %ld\n", cases, bad);
    return bad != 0;
}
