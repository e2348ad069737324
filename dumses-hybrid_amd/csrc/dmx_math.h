// dmx_math.h -- per-cell / per-face / per-edge arithmetic of the MHD step.
//
// Device functions templated on the bulk real type T (double or float).  In
// double they evaluate every expression in the reference's order (Fortran
// left-to-right association; MAX(a,b) = a > b ? a : b, MIN(a,b) = a < b ? a : b,
// SIGN(1,x) = copysign(1,x)) so that the float64 build, compiled with
// -ffp-contract=off, reproduces the reference bit for bit.  Each function
// names the reference routine it computes (paths relative to
// marcjoos-cea/dumses-hybrid/src).
#pragma once

#include <hip/hip_runtime.h>

namespace dmx {

template <typename T> struct K {
    static constexpr T zero = T(0), one = T(1), two = T(2), half = T(0.5), forth = T(0.25);
};
// smallr = smallc = 1e-10 (src/modules/commons.f90:50-51)
__device__ __forceinline__ double smallr_d() { return 1e-10; }

template <typename T> __device__ __forceinline__ T fmx(T a, T b) { return a > b ? a : b; }
template <typename T> __device__ __forceinline__ T fmn(T a, T b) { return a < b ? a : b; }
// float bulk (no bit-exactness to keep): one v_max_f32 / v_min_f32; differs
// from the comparison form only for NaN and the sign of a zero result
template <> __device__ __forceinline__ float fmx<float>(float a, float b) { return __builtin_fmaxf(a, b); }
template <> __device__ __forceinline__ float fmn<float>(float a, float b) { return __builtin_fminf(a, b); }
template <typename T> __device__ __forceinline__ T sgn1(T x) { return copysign(T(1), x); }
template <typename T> __device__ __forceinline__ T fab(T x) { return fabs(x); }

// TVD slope (minmod st=1, MC st=2), scheme.f90:452-477
template <typename T>
__device__ __forceinline__ T tvd_slope(T qm1, T q0, T qp1, int st)
{
    const T s = T(st);
    T dlft = s * (q0 - qm1);
    T drgt = s * (qp1 - q0);
    T dcen = K<T>::half * (dlft + drgt) / s;
    T dsgn = sgn1(dcen);
    T slop = fmn(fab(dlft), fab(drgt));
    T dlim = slop;
    if ((dlft * drgt) <= K<T>::zero) dlim = K<T>::zero;
    return dsgn * fmn(dlim, fab(dcen));
}

// fast magnetosonic speed, the expression shared by comp_speed
// (physics.f90:148-164) and the Riemann solvers (solver.f90:675-683):
// b2 is passed already summed in the caller's order
template <typename T>
__device__ __forceinline__ T cfast_b2(T gamma, T ploc, T r, T bn, T b2)
{
    T c2 = gamma * ploc / r;
    T d2 = K<T>::half * (c2 + b2 / r);
    return sqrt(d2 + sqrt(d2 * d2 - c2 * bn * bn / r));
}

// ---------------------------------------------------------------------------
// 1D Riemann solvers, solver.f90:393-935.
// q = [r, p, u_n, B_n, u_t1, B_t1, u_t2, B_t2]; fg[8]; pstar = total pressure
// of the sampled state (HLLD only; 0 for the others, as pgstar is zeroed).
//
// The energy chain -- total energies and pressures, v.B, the star-state
// energies, the energy flux, and the total-pressure differences that set
// the contact speed -- is evaluated in float64 (type D) in both precisions.
// At low beta the magnetic energy dominates each of these terms and their
// differences carry the thermal energy flux, which float32 would resolve
// only to about eps_32 / beta.  In the float64 build D is T and every
// expression keeps the reference's order, so the bits are unchanged.
// ---------------------------------------------------------------------------
typedef double D;

template <typename T>
__device__ __forceinline__ void phys_flux(T r, T ploc, T u, T v, T w, T a, T b, T c, D entho,
                                          T* f, T* cons, D& fE, D& cE)
{
    const T half = K<T>::half;
    T emag = half * ((a * a + b * b) + c * c);
    T ptot = ploc + emag;
    // energy chain in float64
    D ecinD = 0.5 * ((D(u) * u + D(v) * v) + D(w) * w) * r;
    D emagD = 0.5 * ((D(a) * a + D(b) * b) + D(c) * c);
    D etotD = (D(ploc) * entho + ecinD) + emagD;
    D ptotD = D(ploc) + emagD;
    fE = (etotD + ptotD) * u - D(a) * ((D(u) * a + D(v) * b) + D(w) * c);
    cE = etotD;
    f[0] = r * u;
    f[1] = T(fE);
    f[2] = (r * u * u + ptot) - a * a;
    f[3] = K<T>::zero;
    f[4] = r * u * v - a * b;
    f[5] = b * u - a * v;
    f[6] = r * u * w - a * c;
    f[7] = c * u - a * w;
    cons[0] = r; cons[1] = T(etotD); cons[2] = r * u; cons[3] = a;
    cons[4] = r * v; cons[5] = b; cons[6] = r * w; cons[7] = c;
}

// hlld, solver.f90:618-836
template <typename T>
__device__ __forceinline__ void hlld1d(const T* ql, const T* qr, T gamma, T ploc, T proc,
                                       T* fg, T& pstar)
{
    const T zero = K<T>::zero, half = K<T>::half;
    T rl = ql[0], ul = ql[2], al = ql[3], vl = ql[4], bl = ql[5], wl = ql[6], cl = ql[7];
    T rr = qr[0], ur = qr[2], ar = qr[3], vr = qr[4], br = qr[5], wr = qr[6], cr = qr[7];
    const D entho = 1.0 / (D(gamma) - 1.0);
    T a = half * (al + ar);
    T sgnm = sgn1(a);
    D ecinl = 0.5 * ((D(ul) * ul + D(vl) * vl) + D(wl) * wl) * rl;
    D emagl = 0.5 * ((D(a) * a + D(bl) * bl) + D(cl) * cl);
    D etotl = (D(ploc) * entho + ecinl) + emagl;
    D ptotl = D(ploc) + emagl;
    D vdotbl = (D(ul) * a + D(vl) * bl) + D(wl) * cl;
    D ecinr = 0.5 * ((D(ur) * ur + D(vr) * vr) + D(wr) * wr) * rr;
    D emagr = 0.5 * ((D(a) * a + D(br) * br) + D(cr) * cr);
    D etotr = (D(proc) * entho + ecinr) + emagr;
    D ptotr = D(proc) + emagr;
    D vdotbr = (D(ur) * a + D(vr) * br) + D(wr) * cr;
    T cfastl = cfast_b2(gamma, ploc, rl, a, (a * a + bl * bl) + cl * cl);
    T cfastr = cfast_b2(gamma, proc, rr, a, (a * a + br * br) + cr * cr);
    T sl = fmn(ul, ur) - fmx(cfastl, cfastr);
    T sr = fmx(ul, ur) + fmx(cfastl, cfastr);
    T rcl = rl * (ul - sl);
    T rcr = rr * (sr - ur);
    const D ustarD = ((D(rcr) * ur + D(rcl) * ul) + (ptotl - ptotr)) / (D(rcr) + rcl);
    const T ustar = T(ustarD);
    const D ptotstar = ((D(rcr) * ptotl + D(rcl) * ptotr) + D(rcl) * rcr * (D(ul) - ur)) / (D(rcr) + rcl);
    T vstarl, bstarl, wstarl, cstarl, vstarr, bstarr, wstarr, cstarr;
    T rstarl = rl * (sl - ul) / (sl - ustar);
    T estar = rl * (sl - ul) * (sl - ustar) - a * a;
    T el = rl * (sl - ul) * (sl - ul) - a * a;
    if (estar == zero) {
        vstarl = vl; bstarl = bl; wstarl = wl; cstarl = cl;
    } else {
        vstarl = vl - a * bl * (ustar - ul) / estar;
        bstarl = bl * el / estar;
        wstarl = wl - a * cl * (ustar - ul) / estar;
        cstarl = cl * el / estar;
    }
    D vdotbstarl = (ustarD * a + D(vstarl) * bstarl) + D(wstarl) * cstarl;
    D etotstarl = ((((D(sl) - ul) * etotl - ptotl * ul) + ptotstar * ustarD) + D(a) * (vdotbl - vdotbstarl))
                  / (D(sl) - ustarD);
    T sqrrstarl = sqrt(rstarl);
    T calfvenl = fab(a) / sqrrstarl;
    T sal = ustar - calfvenl;
    T rstarr = rr * (sr - ur) / (sr - ustar);
    estar = rr * (sr - ur) * (sr - ustar) - a * a;
    T er = rr * (sr - ur) * (sr - ur) - a * a;
    if (estar == zero) {
        vstarr = vr; bstarr = br; wstarr = wr; cstarr = cr;
    } else {
        vstarr = vr - a * br * (ustar - ur) / estar;
        bstarr = br * er / estar;
        wstarr = wr - a * cr * (ustar - ur) / estar;
        cstarr = cr * er / estar;
    }
    D vdotbstarr = (ustarD * a + D(vstarr) * bstarr) + D(wstarr) * cstarr;
    D etotstarr = ((((D(sr) - ur) * etotr - ptotr * ur) + ptotstar * ustarD) + D(a) * (vdotbr - vdotbstarr))
                  / (D(sr) - ustarD);
    T sqrrstarr = sqrt(rstarr);
    T calfvenr = fab(a) / sqrrstarr;
    T sar = ustar + calfvenr;
    T den = sqrrstarl + sqrrstarr;
    T vss = ((sqrrstarl * vstarl + sqrrstarr * vstarr) + sgnm * (bstarr - bstarl)) / den;
    T wss = ((sqrrstarl * wstarl + sqrrstarr * wstarr) + sgnm * (cstarr - cstarl)) / den;
    T bss = ((sqrrstarl * bstarr + sqrrstarr * bstarl) + sgnm * sqrrstarl * sqrrstarr * (vstarr - vstarl)) / den;
    T css = ((sqrrstarl * cstarr + sqrrstarr * cstarl) + sgnm * sqrrstarl * sqrrstarr * (wstarr - wstarl)) / den;
    D vdotbss = (ustarD * a + D(vss) * bss) + D(wss) * css;
    D etotssl = etotstarl - D(sgnm) * sqrrstarl * (vdotbstarl - vdotbss);
    D etotssr = etotstarr + D(sgnm) * sqrrstarr * (vdotbstarr - vdotbss);
    T ro, uo, vo, wo, bo, co;
    D uoD, ptoto, etoto, vdotbo;
    if (sl > zero) {
        ro = rl; uo = ul; vo = vl; wo = wl; bo = bl; co = cl; uoD = ul; ptoto = ptotl; etoto = etotl; vdotbo = vdotbl;
    } else if (sal > zero) {
        ro = rstarl; uo = ustar; vo = vstarl; wo = wstarl; bo = bstarl; co = cstarl; uoD = ustarD; ptoto = ptotstar;
        etoto = etotstarl; vdotbo = vdotbstarl;
    } else if (ustar > zero) {
        ro = rstarl; uo = ustar; vo = vss; wo = wss; bo = bss; co = css; uoD = ustarD; ptoto = ptotstar;
        etoto = etotssl; vdotbo = vdotbss;
    } else if (sar > zero) {
        ro = rstarr; uo = ustar; vo = vss; wo = wss; bo = bss; co = css; uoD = ustarD; ptoto = ptotstar;
        etoto = etotssr; vdotbo = vdotbss;
    } else if (sr > zero) {
        ro = rstarr; uo = ustar; vo = vstarr; wo = wstarr; bo = bstarr; co = cstarr; uoD = ustarD; ptoto = ptotstar;
        etoto = etotstarr; vdotbo = vdotbstarr;
    } else {
        ro = rr; uo = ur; vo = vr; wo = wr; bo = br; co = cr; uoD = ur; ptoto = ptotr; etoto = etotr; vdotbo = vdotbr;
    }
    fg[0] = ro * uo;
    fg[1] = T((etoto + ptoto) * uoD - D(a) * vdotbo);
    fg[2] = ro * uo * uo - a * a;
    fg[3] = zero;
    fg[4] = ro * uo * vo - a * bo;
    fg[5] = bo * uo - a * vo;
    fg[6] = ro * uo * wo - a * co;
    fg[7] = co * uo - a * wo;
    pstar = T(ptoto);
}

// llf (solver.f90:393-498), hll (:504-612), upwind (:842-935)
template <typename T, int SOLVER>
__device__ __forceinline__ void simple1d(const T* ql, const T* qr, T gamma, T ploc, T proc, T* fg)
{
    const T zero = K<T>::zero, half = K<T>::half;
    T rl = ql[0], ul = ql[2], al = ql[3], vl = ql[4], bl = ql[5], wl = ql[6], cl = ql[7];
    T rr = qr[0], ur = qr[2], ar = qr[3], vr = qr[4], br = qr[5], wr = qr[6], cr = qr[7];
    const D entho = 1.0 / (D(gamma) - 1.0);
    T a = half * (al + ar);
    T fl[8], cL[8], fr[8], cR[8];
    D flE, cLE, frE, cRE;
    phys_flux(rl, ploc, ul, vl, wl, a, bl, cl, entho, fl, cL, flE, cLE);
    phys_flux(rr, proc, ur, vr, wr, a, br, cr, entho, fr, cR, frE, cRE);
    if (SOLVER == 3) { // hll
        T vleft = cfast_b2(gamma, ploc, rl, a, (a * a + bl * bl) + cl * cl);
        T vright = cfast_b2(gamma, proc, rr, a, (a * a + br * br) + cr * cr);
        T sl = fmn(fmn(ul, ur) - fmx(vleft, vright), zero);
        T sr = fmx(fmx(ul, ur) + fmx(vleft, vright), zero);
#pragma unroll
        for (int n = 0; n < 8; n++) fg[n] = ((sr * fl[n] - sl * fr[n]) + sr * sl * (cR[n] - cL[n])) / (sr - sl);
        fg[1] = T(((D(sr) * flE - D(sl) * frE) + D(sr) * sl * (cRE - cLE)) / (D(sr) - sl));
    } else {
        T vel;
        if (SOLVER == 2) { // llf
            T vleft = cfast_b2(gamma, ploc, rl, a, (a * a + bl * bl) + cl * cl) + fab(ul);
            T vright = cfast_b2(gamma, proc, rr, a, (a * a + br * br) + cr * cr) + fab(ur);
            vel = fmx(vleft, vright);
        } else { // upwind
            vel = fab(half * (ul + ur));
        }
#pragma unroll
        for (int n = 0; n < 8; n++) fg[n] = half * (fl[n] + fr[n]) - vel * half * (cR[n] - cL[n]);
        fg[1] = T(0.5 * (flE + frE) - D(vel) * 0.5 * (cRE - cLE));
    }
}

template <typename T, int SOLVER>
__device__ __forceinline__ void riemann1d(const T* ql, const T* qr, T gamma, T c2iso, bool iso,
                                          T* fg, T& pstar)
{
    T ploc = iso ? ql[0] * c2iso : ql[1];
    T proc = iso ? qr[0] * c2iso : qr[1];
    if (SOLVER == 4) {
        hlld1d(ql, qr, gamma, ploc, proc, fg, pstar);
    } else {
        simple1d<T, SOLVER>(ql, qr, gamma, ploc, proc, fg);
        pstar = K<T>::zero;
    }
}

// ---------------------------------------------------------------------------
// 2D (edge) Riemann solvers, solver_magnetic.f90:241-727.
// corner q = [r, p, u, v, A, B, C, E]; returns emf*dt
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ T cf2x(T gamma, const T* q)
{
    return cfast_b2(gamma, q[1], q[0], q[4], (q[4] * q[4] + q[5] * q[5]) + q[6] * q[6]);
}
template <typename T>
__device__ __forceinline__ T cf2y(T gamma, const T* q)
{
    return cfast_b2(gamma, q[1], q[0], q[5], (q[5] * q[5] + q[4] * q[4]) + q[6] * q[6]);
}

// hlld, solver_magnetic.f90:362-548
template <typename T>
__device__ __forceinline__ T hlld2d(const T* LL, const T* RL, const T* LR, const T* RR, T dt, T gamma)
{
    const T zero = K<T>::zero, one = K<T>::one, half = K<T>::half;
    const T smallc = T(1e-10);
    T rLL = LL[0], pLL = LL[1], uLL = LL[2], vLL = LL[3], ALL = LL[4], BLL = LL[5], CLL = LL[6], ELL = LL[7];
    T rRL = RL[0], pRL = RL[1], uRL = RL[2], vRL = RL[3], ARL = RL[4], BRL = RL[5], CRL = RL[6], ERL = RL[7];
    T rLR = LR[0], pLR = LR[1], uLR = LR[2], vLR = LR[3], ALR = LR[4], BLR = LR[5], CLR = LR[6], ELR = LR[7];
    T rRR = RR[0], pRR = RR[1], uRR = RR[2], vRR = RR[3], ARR = RR[4], BRR = RR[5], CRR = RR[6], ERR = RR[7];
    T cLLx = cf2x(gamma, LL), cLRx = cf2x(gamma, LR), cRLx = cf2x(gamma, RL), cRRx = cf2x(gamma, RR);
    T cLLy = cf2y(gamma, LL), cLRy = cf2y(gamma, LR), cRLy = cf2y(gamma, RL), cRRy = cf2y(gamma, RR);
    T SL = fmn(fmn(fmn(uLL, uLR), uRL), uRR) - fmx(fmx(fmx(cLLx, cLRx), cRLx), cRRx);
    T SR = fmx(fmx(fmx(uLL, uLR), uRL), uRR) + fmx(fmx(fmx(cLLx, cLRx), cRLx), cRRx);
    T SB = fmn(fmn(fmn(vLL, vLR), vRL), vRR) - fmx(fmx(fmx(cLLy, cLRy), cRLy), cRRy);
    T ST = fmx(fmx(fmx(vLL, vLR), vRL), vRR) + fmx(fmx(fmx(cLLy, cLRy), cRLy), cRRy);
    // total pressures in float64 (D): their differences set the contact
    // speeds, and at low beta the magnetic pressure dominates each of them
    D PtotLL = D(pLL) + 0.5 * ((D(ALL) * ALL + D(BLL) * BLL) + D(CLL) * CLL);
    D PtotLR = D(pLR) + 0.5 * ((D(ALR) * ALR + D(BLR) * BLR) + D(CLR) * CLR);
    D PtotRL = D(pRL) + 0.5 * ((D(ARL) * ARL + D(BRL) * BRL) + D(CRL) * CRL);
    D PtotRR = D(pRR) + 0.5 * ((D(ARR) * ARR + D(BRR) * BRR) + D(CRR) * CRR);
    T rcLLx = rLL * (uLL - SL), rcRLx = rRL * (SR - uRL);
    T rcLRx = rLR * (uLR - SL), rcRRx = rRR * (SR - uRR);
    T rcLLy = rLL * (vLL - SB), rcLRy = rLR * (ST - vLR);
    T rcRLy = rRL * (vRL - SB), rcRRy = rRR * (ST - vRR);
    T ustar = T(((((D(rcLLx) * uLL + D(rcLRx) * uLR) + D(rcRLx) * uRL) + D(rcRRx) * uRR)
                 + (((PtotLL - PtotRL) + PtotLR) - PtotRR)) / (((D(rcLLx) + rcLRx) + rcRLx) + rcRRx));
    T vstar = T(((((D(rcLLy) * vLL + D(rcLRy) * vLR) + D(rcRLy) * vRL) + D(rcRRy) * vRR)
                 + (((PtotLL - PtotLR) + PtotRL) - PtotRR)) / (((D(rcLLy) + rcLRy) + rcRLy) + rcRRy));
    T SLustar = one / (SL - ustar), SRustar = one / (SR - ustar);
    T SBvstar = one / (SB - vstar), STvstar = one / (ST - vstar);

    T rstarLLx = rLL * (SL - uLL) * SLustar, BstarLL = BLL * (SL - uLL) * SLustar;
    T rstarLLy = rLL * (SB - vLL) * SBvstar, AstarLL = ALL * (SB - vLL) * SBvstar;
    T rstarLL = rLL * (SL - uLL) * SLustar * (SB - vLL) * SBvstar;
    T EstarLLx = ustar * BstarLL - vLL * ALL, EstarLLy = uLL * BLL - vstar * AstarLL;
    T EstarLL = ustar * BstarLL - vstar * AstarLL;

    T rstarLRx = rLR * (SL - uLR) * SLustar, BstarLR = BLR * (SL - uLR) * SLustar;
    T rstarLRy = rLR * (ST - vLR) * STvstar, AstarLR = ALR * (ST - vLR) * STvstar;
    T rstarLR = rLR * (SL - uLR) * SLustar * (ST - vLR) * STvstar;
    T EstarLRx = ustar * BstarLR - vLR * ALR, EstarLRy = uLR * BLR - vstar * AstarLR;
    T EstarLR = ustar * BstarLR - vstar * AstarLR;

    T rstarRLx = rRL * (SR - uRL) * SRustar, BstarRL = BRL * (SR - uRL) * SRustar;
    T rstarRLy = rRL * (SB - vRL) * SBvstar, AstarRL = ARL * (SB - vRL) * SBvstar;
    T rstarRL = rRL * (SR - uRL) * SRustar * (SB - vRL) * SBvstar;
    T EstarRLx = ustar * BstarRL - vRL * ARL, EstarRLy = uRL * BRL - vstar * AstarRL;
    T EstarRL = ustar * BstarRL - vstar * AstarRL;

    T rstarRRx = rRR * (SR - uRR) * SRustar, BstarRR = BRR * (SR - uRR) * SRustar;
    T rstarRRy = rRR * (ST - vRR) * STvstar, AstarRR = ARR * (ST - vRR) * STvstar;
    T rstarRR = rRR * (SR - uRR) * SRustar * (ST - vRR) * STvstar;
    T EstarRRx = ustar * BstarRR - vRR * ARR, EstarRRy = uRR * BRR - vstar * AstarRR;
    T EstarRR = ustar * BstarRR - vstar * AstarRR;

    T sqLR = one / sqrt(rstarLR), sqRL = one / sqrt(rstarRL);
    T sqRR = one / sqrt(rstarRR), sqLL = one / sqrt(rstarLL);

    T calfvenL = fmx(fmx(fmx(fmx(fab(ALR) / sqrt(rstarLRx), fab(AstarLR) * sqLR),
                             fab(ALL) / sqrt(rstarLLx)), fab(AstarLL) * sqLL), smallc);
    T calfvenR = fmx(fmx(fmx(fmx(fab(ARR) / sqrt(rstarRRx), fab(AstarRR) * sqRR),
                             fab(ARL) / sqrt(rstarRLx)), fab(AstarRL) * sqRL), smallc);
    T calfvenB = fmx(fmx(fmx(fmx(fab(BLL) / sqrt(rstarLLy), fab(BstarLL) * sqLL),
                             fab(BRL) / sqrt(rstarRLy)), fab(BstarRL) * sqRL), smallc);
    T calfvenT = fmx(fmx(fmx(fmx(fab(BLR) / sqrt(rstarLRy), fab(BstarLR) * sqLR),
                             fab(BRR) / sqrt(rstarRRy)), fab(BstarRR) * sqRR), smallc);
    T SAL = fmn(ustar - calfvenL, zero), SAR = fmx(ustar + calfvenR, zero);
    T SAB = fmn(vstar - calfvenB, zero), SAT = fmx(vstar + calfvenT, zero);
    T SARSAL = one / (SAR - SAL);
    T SATSAB = one / (SAT - SAB);
    T AstarT = (SAR * AstarRR - SAL * AstarLR) * SARSAL;
    T AstarB = (SAR * AstarRL - SAL * AstarLL) * SARSAL;
    T BstarR = (SAT * BstarRR - SAB * BstarRL) * SATSAB;
    T BstarL = (SAT * BstarLR - SAB * BstarLL) * SATSAB;
    T E;
    if (SB > zero) {
        if (SL > zero) E = ELL;
        else if (SR < zero) E = ERL;
        else E = ((SAR * EstarLLx - SAL * EstarRLx) + SAR * SAL * (BRL - BLL)) * SARSAL;
    } else if (ST < zero) {
        if (SL > zero) E = ELR;
        else if (SR < zero) E = ERR;
        else E = ((SAR * EstarLRx - SAL * EstarRRx) + SAR * SAL * (BRR - BLR)) * SARSAL;
    } else if (SL > zero) {
        E = ((SAT * EstarLLy - SAB * EstarLRy) - SAT * SAB * (ALR - ALL)) * SATSAB;
    } else if (SR < zero) {
        E = ((SAT * EstarRLy - SAB * EstarRRy) - SAT * SAB * (ARR - ARL)) * SATSAB;
    } else {
        E = (((((SAL * SAB * EstarRR - SAL * SAT * EstarRL) - SAR * SAB * EstarLR) + SAR * SAT * EstarLL) * SARSAL) * SATSAB
             - SAT * SAB * SATSAB * (AstarT - AstarB))
            + SAR * SAL * SARSAL * (BstarR - BstarL);
    }
    return E * dt;
}

template <typename T>
__device__ __forceinline__ T llf2d_dir(T gamma, T rl, T pl, T ul, T al, T bl, T cl, T rr, T pr, T ur,
                                       T ar, T br, T cr, bool iso, T c2iso)
{
    // one direction of llf, solver_magnetic.f90:267-297
    const T half = K<T>::half;
    if (iso) {
        pl = rl * c2iso;
        pr = rr * c2iso;
    }
    al = half * (al + ar);
    ar = al;
    T vleft = cfast_b2(gamma, pl, rl, al, (al * al + bl * bl) + cl * cl) + fab(ul);
    T vright = cfast_b2(gamma, pr, rr, ar, (ar * ar + br * br) + cr * cr) + fab(ur);
    T vel = fmx(vleft, vright);
    return -vel * half * (br - bl);
}

template <typename T, int SOLVER>
__device__ __forceinline__ T riemann2d(const T* LL, const T* RL, const T* LR, const T* RR, T dt,
                                       T gamma, bool iso, T c2iso)
{
    const T zero = K<T>::zero, half = K<T>::half, forth = K<T>::forth;
    if (SOLVER == 4) return hlld2d(LL, RL, LR, RR, dt, gamma);
    if (SOLVER == 2) { // llf, solver_magnetic.f90:241-356
        T E = forth * (((LL[7] + RL[7]) + LR[7]) + RR[7]);
        T fx = llf2d_dir(gamma, half * (LL[0] + LR[0]), half * (LL[1] + LR[1]), half * (LL[2] + LR[2]),
                         half * (LL[4] + LR[4]), half * (LL[5] + LR[5]), half * (LL[6] + LR[6]),
                         half * (RR[0] + RL[0]), half * (RR[1] + RL[1]), half * (RR[2] + RL[2]),
                         half * (RR[4] + RL[4]), half * (RR[5] + RL[5]), half * (RR[6] + RL[6]), iso, c2iso);
        T fy = llf2d_dir(gamma, half * (LL[0] + RL[0]), half * (LL[1] + RL[1]), half * (LL[3] + RL[3]),
                         half * (LL[5] + RL[5]), half * (LL[4] + RL[4]), half * (LL[6] + RL[6]),
                         half * (RR[0] + LR[0]), half * (RR[1] + LR[1]), half * (RR[3] + LR[3]),
                         half * (RR[5] + LR[5]), half * (RR[4] + LR[4]), half * (RR[6] + LR[6]), iso, c2iso);
        T e = E + (fx - fy);
        return e * dt;
    }
    if (SOLVER == 5) { // upwind, solver_magnetic.f90:554-600
        T E = forth * (((LL[7] + RL[7]) + LR[7]) + RR[7]);
        T ul = half * (LL[2] + LR[2]), bl = half * (LL[5] + LR[5]);
        T ur = half * (RR[2] + RL[2]), br = half * (RR[5] + RL[5]);
        T vel = half * (ul + ur);
        T fx = -fab(vel) * half * (br - bl);
        ul = half * (LL[3] + RL[3]); bl = half * (LL[4] + RL[4]);
        ur = half * (RR[3] + LR[3]); br = half * (RR[4] + LR[4]);
        vel = half * (ul + ur);
        T fy = -fab(vel) * half * (br - bl);
        T e = E + (fx - fy);
        return e * dt;
    }
    // hlla, solver_magnetic.f90:688-727
    const T smallc = T(1e-10);
    T cLLx = sqrt(LL[4] * LL[4] / LL[0]), cLRx = sqrt(LR[4] * LR[4] / LR[0]);
    T cRLx = sqrt(RL[4] * RL[4] / RL[0]), cRRx = sqrt(RR[4] * RR[4] / RR[0]);
    T cLLy = sqrt(LL[5] * LL[5] / LL[0]), cLRy = sqrt(LR[5] * LR[5] / LR[0]);
    T cRLy = sqrt(RL[5] * RL[5] / RL[0]), cRRy = sqrt(RR[5] * RR[5] / RR[0]);
    T SL = fmn(fmn(fmn(fmn(LL[2], LR[2]), RL[2]), RR[2]) - fmx(fmx(fmx(fmx(cLLx, cLRx), cRLx), cRRx), smallc), zero);
    T SR = fmx(fmx(fmx(fmx(LL[2], LR[2]), RL[2]), RR[2]) + fmx(fmx(fmx(fmx(cLLx, cLRx), cRLx), cRRx), smallc), zero);
    T SB = fmn(fmn(fmn(fmn(LL[3], LR[3]), RL[3]), RR[3]) - fmx(fmx(fmx(fmx(cLLy, cLRy), cRLy), cRRy), smallc), zero);
    T ST = fmx(fmx(fmx(fmx(LL[3], LR[3]), RL[3]), RR[3]) + fmx(fmx(fmx(fmx(cLLy, cLRy), cRLy), cRRy), smallc), zero);
    T e = ((((SL * SB * RR[7] - SL * ST * RL[7]) - SR * SB * LR[7]) + SR * ST * LL[7]) / (SR - SL) / (ST - SB)
           - ST * SB / (ST - SB) * (RR[4] - LL[4]))
          + SR * SL / (SR - SL) * (RR[5] - LL[5]);
    return e * dt;
}

} // namespace dmx
