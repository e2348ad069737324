// dmx_tiling.h -- how the z-marching kernels cut a domain into column
// tiles and blocks of planes.  Plain integer code for host and device;
// tests/tiling_check.cpp compiles it on the host and checks that the tiles
// cover every column and plane exactly once within their limits.
#pragma once

#ifndef __HIPCC__
#ifndef __host__
#define __host__
#endif
#ifndef __device__
#define __device__
#endif
#ifndef __forceinline__
#define __forceinline__ inline
#endif
#endif

namespace dmx {

// A run of n columns cut into the fewest tiles of at most wmax: nt tiles,
// the first nt - nwide of them w wide, the last nwide w + 1 wide.  With
// wmax one more than a power-of-two tile width, the 2^m + 1 faces or
// 2^m + 2 cells of a power-of-two domain need no tile of one or two
// columns, which would cost a z-marching block as much as a full one.
struct Part { int nt, w, nwide; };
__host__ __device__ __forceinline__ Part make_part(int n, int wmax)
{
    Part p;
    p.nt = n > 0 ? (n + wmax - 1) / wmax : 1;
    p.w = n / p.nt;
    p.nwide = n - p.w * p.nt;
    return p;
}
__host__ __device__ __forceinline__ void part_tile(const Part& p, int t, int& start, int& width)
{
    const int nn = p.nt - p.nwide;
    start = t * p.w + (t > nn ? t - nn : 0);
    width = p.w + (t >= nn ? 1 : 0);
}

// The columns [0, nxc) x [0, nyc) of a plane in tiles of at most fc
// columns, at most fxm wide and fym high: x strips of the fewest tiles
// (make_part), each cut into rows of at most min(fym, fc / width).  Tiles
// are numbered strip by strip, rows fastest; nz blocks of planes, the
// fewest of at least kt planes (between kt and 2 kt - 1), follow one
// another.
struct Tiling {
    Part px, pyn, pyw, pz;
    int tz;   // tiles of one block of planes
};
__host__ __device__ __forceinline__ Tiling make_tiling(int nxc, int nyc, int nzc, int fxm, int fym, int fc, int kt)
{
    Tiling t;
    t.px = make_part(nxc, fxm);
    const int wn = t.px.w > 0 ? t.px.w : 1, ww = wn + 1;
    t.pyn = make_part(nyc, fc / wn < fym ? fc / wn : fym);
    t.pyw = make_part(nyc, fc / ww < fym ? fc / ww : fym);
    t.tz = (t.px.nt - t.px.nwide) * t.pyn.nt + t.px.nwide * t.pyw.nt;
    const int nzb = nzc / kt > 0 ? nzc / kt : 1;
    t.pz = make_part(nzc, (nzc + nzb - 1) / nzb);
    return t;
}
// tile l of block of planes zb: its first column, row and plane (from 0)
// and its extents
__host__ __device__ __forceinline__ void tiling_tile(const Tiling& t, int l, int zb, int& x0, int& fx, int& y0, int& fy,
                                                     int& z0, int& fz)
{
    const int nn = t.px.nt - t.px.nwide;
    int xt, yt;
    if (l < nn * t.pyn.nt) {
        xt = l / t.pyn.nt;
        yt = l - xt * t.pyn.nt;
        part_tile(t.pyn, yt, y0, fy);
    } else {
        l -= nn * t.pyn.nt;
        xt = nn + l / t.pyw.nt;
        yt = l - (xt - nn) * t.pyw.nt;
        part_tile(t.pyw, yt, y0, fy);
    }
    part_tile(t.px, xt, x0, fx);
    part_tile(t.pz, zb, z0, fz);
}

} // namespace dmx
