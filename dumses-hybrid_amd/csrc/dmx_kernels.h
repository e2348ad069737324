// dmx_kernels.h -- device data layout and kernels of the MHD step.
//
// Layout in HBM (one context = one ghosted domain, nghost = 3):
//   index(i,j,k) = (i+2) + px*((j+2) + gy*(k+2)),  i in [-2, nx+3] (the
//   reference's uin index range), px = nx+6 rounded up to 16 elements so
//   every x-row starts 128-byte aligned; x is the fastest index (coalesced).
//   U  : 8 planes -- rho, rho*u, rho*v, rho*w in the bulk type T; the total
//        energy E and the left-face fields Bx_L, By_L, Bz_L always in float64
//        (E: the pressure recovery at low beta subtracts the magnetic energy
//        from it; B: the CT face-B accumulation stays float64 in both
//        precisions).  The reference's right faces are implied:
//        B_R(i) = B_L(i+1).
//   S  : 47 values of T per cell, the per-cell trace summary written by the
//        trace kernel for cells [0, n+1]^3 (see TraceSlot), in groups of
//        256 bytes of consecutive cells per slot (s_off in dmx_step.inc).
//   F  : 18 planes of T, x/y/z face fluxes of rho, rho*u, rho*v, rho*w, E
//        (already multiplied by ds*dt as compflux does) and the HLLD star
//        total pressure p_T (raw).
//   EM : 3 planes of T, edge EMFs times dt.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "dmx_tiling.h"

namespace dmx {

enum TraceSlot {
    S_R = 0, S_U, S_V, S_W, S_P, S_A, S_B, S_C,          // predicted centre state
    S_DRX = 8, S_DUX, S_DVX, S_DWX, S_DPX, S_DBX, S_DCX, // half slopes along x
    S_DRY = 15, S_DUY, S_DVY, S_DWY, S_DPY, S_DAY, S_DCY,
    S_DRZ = 22, S_DUZ, S_DVZ, S_DWZ, S_DPZ, S_DAZ, S_DBZ,
    S_AL = 29, S_AR, S_BL, S_BR, S_CL, S_CR,             // predicted face fields
    S_DALY = 35, S_DARY, S_DALZ, S_DARZ,                 // face-field half slopes
    S_DBLX, S_DBRX, S_DBLZ, S_DBRZ,
    S_DCLX, S_DCRX, S_DCLY, S_DCRY,
    NSLOT = 47
};

// flux planes: F[d*6 + c], c in {rho, mx, my, mz, E, pstar}
enum { F_R = 0, F_MX, F_MY, F_MZ, F_E, F_PS, NFC = 6 };

struct Geom {
    int nx, ny, nz;
    int px, gy, gz;
    long long nvol;      // elements per plane
    double dx, dy, dz;
    double xmin, xmax, ymin, ymax, zmin, zmax;
    __host__ __device__ __forceinline__ long long idx(int i, int j, int k) const
    {
        return (long long)(i + 2) + (long long)px * ((long long)(j + 2) + (long long)gy * (k + 2));
    }
    __device__ __forceinline__ double xc(int i) const { return xmin + dx / 2 + (double)(i - 1) * dx; }
};

struct Phys {
    double gamma, ciso, omega0, courant;
    double nu, eta;
    int iso, slope_type, riemann, riemann2d, fargo;
    int bc[3];
};

template <typename T> struct Fields {
    T* uh;          // 4 planes: rho, rho*u, rho*v, rho*w
    double* ue;     // total energy E
    double* ub;     // 3 face-B planes
    T* s;           // NSLOT planes
    T* f;           // 18 planes
    T* em;          // 3 planes
    // learned limiter (slope_type 4) only
    float* lq;      // 8 planes: primitive variables of every cell
    float* ls;      // 24 planes: limited slopes [direction][variable]
    const unsigned short* lw;   // packed bf16 weights (see the limiter kernel)
    const float* lb;            // biases b1 (64), b2 (64), b3 (8)
    unsigned int* lcount;       // fallback counter
};

} // namespace dmx
