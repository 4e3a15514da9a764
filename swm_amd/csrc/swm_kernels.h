// Device kernels of the SWM time-stepping path for gfx950 (CDNA4).
//
// One fused kernel advances u, v, eta and up to two tracers by one time step:
// it reproduces SWM_timestep_step + TRC_timestep + model_advance of the
// reference (src/swm_timestep_module.f90:164-180,328-449,
// src/swm_lateralmixing_module.f90:157-240, src/tracer_module.f90:154-280,
// src/time_integration_module.f90:74-112, src/swm_module.f90:86-139).
// Every field the reference materialises between loops (D, Dh, Du, Dv, zeta,
// Pot, EDens, MU, MV, pll, plt, latmix, F_x/F_y/F_eta, C) is recomputed from
// the state staged in LDS, so HBM sees only the state, its AB history and the
// packed mask.
//
// Arithmetic mirrors the reference expression trees operation by operation
// (left-to-right sums, divisions kept as divisions, no contraction: the file
// is compiled with -ffp-contract=off), which makes the FP64 path bit-identical
// to the reference-order CPU restatement.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace swm {

enum : uint8_t { OC_H = 1, OC_U = 2, OC_V = 4, OC_E = 8 };
enum { VAR_FN = 1, VAR_SOR = 3 };

// A static per-cell input: either a full (Ny x P) array or, when every row holds
// one value on the grid's ocean points and one on its land points, two values
// per row (ocean, land).  The host picks the exact compressed form.
template <typename T>
struct Fld {
    const T* full;
    const T* rows;
};

template <typename T>
__device__ __forceinline__ T fld(const Fld<T>& f, size_t off, int j, bool ocean) {
    return f.full ? f.full[off] : f.rows[2 * j + (ocean ? 0 : 1)];
}

// Static inputs of the step kernel, resident in device memory and read through
// one pointer (scalar loads on use), so the per-launch argument block stays
// small and the kernel does not spill uniform values out of SGPRs.
template <typename T>
struct Statics {
    int Nx, Ny, P;
    int mix, quad, barotropic, has_tau_scale;
    int has_taux, has_tauy, has_ox, has_oy, has_oe;
    T G, A, dL, dT, RHO0, minD, lbc, tau_scale, AdT, eps;
    const uint8_t* mask;
    Fld<T> He, Hu, Hv, imu, imv, ime, gsu, gsv, Fxc, Fyc, Fec, taux, tauy, oRx, oIx, oRy, oIy, oRe, oIe;
    // per-row tables
    const T *cu, *cv, *ce, *cH, *fH;
    const T* za;     // [Ny][2] bc_H/((A*cos_H)*dLambda), ocean_H = 1 / 0
    const T* edu;    // [Ny] (A*cos_u)*dLambda
    const T* ace;    // [Ny] A*cos_eta
    const T* lmu;    // [Ny][2][3] lat_mixing_u at ocean u, o_tmp = 1, 2
    const T* lmv;    // [Ny][2][3]
    const T* pllc;   // [Ny][ov(j+1)][ov(j)][3]
    const T* pltc;   // [Ny][ou(j)][ou(j-1)][3]
    const T *tA, *tB, *tC;   // [Ny][2] tracer advection magnitudes, denominator 1 / 2
    const T *t9, *t11, *t12; // [Ny] tracer diffusion magnitudes
    // tracers
    T kappa[2];
    Fld<T> C0[2], gC[2], timpl[2];
    // diagnostic outputs (materialisation kernel only)
    T* diag[14];      // D Dh Du Dv zeta Pot EDens MU MV F_x F_y F_eta latmix_u latmix_v
    T* tdiag[2][3];   // C uhc vhc per tracer
};

// State planes in one allocation: u, v, eta at two levels, three AB slots per
// field, then per tracer CH (two levels), G_CH (three slots) and the C buffer.
enum : int { PL_U = 0, PL_V = 2, PL_E = 4, PL_GU = 6, PL_GV = 9, PL_GE = 12, PL_TRC = 15, PL_PER_TRC = 6 };

// Per-launch arguments (by value).
template <typename T>
struct Dyn {
    const Statics<T>* s;
    T* state;
    long long plane;     // elements per plane (Ny * P)
    int ntx, nblk;
    int cur, ring;       // ping-pong level of N0; ring slot of G(1)
    int tstep;           // min(3, itt)
    int cbuf;            // bit n: tracer n reads its C buffer (first step after hand-over)
    T dtc1, dtc2, dtc3;  // dt*ab_coeff(ti, tstep)
    T oc[3], os[3];      // cos/sin(omega*itt*dt) of the oscillating streams (x, y, eta)
};

// weight = 1 / max(sum of masks, 1) of the *_noland interpolators
// (src/calc_lib.f90:259-262, 326-332): the exact values of those divisions.
template <typename T>
__device__ __forceinline__ T noland_w(int s) {
    return s <= 1 ? T(1) : (s == 2 ? T(0.5) : (s == 3 ? T(1) / T(3) : T(0.25)));
}

__device__ __forceinline__ int wrap(int x, int n) {
    if (x < 0) x += n;
    if (x >= n) x -= n;
    if (x < 0 || x >= n) x = ((x % n) + n) % n;
    return x;
}

// XCD-aware tile order: the dispatcher sends block b to XCD b % 8; give each
// XCD one contiguous band of tiles so halo rows shared by neighbouring tiles
// are re-read from that XCD's L2.
__device__ __forceinline__ int xcd_tile(int b, int nb) {
    const int q = nb >> 3, r = nb & 7, x = b & 7, k = b >> 3;
    return x * q + (x < r ? x : r) + k;
}

#define OE(m) (((m) >> 3) & 1)
#define OV(m) (((m) >> 2) & 1)
#define OU(m) (((m) >> 1) & 1)
#define OH(m) ((m) & 1)

// interp2 with *_noland weights of D between two eta points (eta2u / eta2v)
template <typename T>
__device__ __forceinline__ T interp_noland2(T D1, T D2, int m1, int m2) {
    const T w = noland_w<T>(m1 + m2);
    return (m1 ? w : T(0)) * D1 + (m2 ? w : T(0)) * D2;
}

// Tracer advection coefficient TRC_coeff(k) (src/tracer_module.f90:318-333):
// sign * num_mask / den / A / d / cos, den = ocean_eta(nb) + ocean_eta(i,j).
// den == 0 is the reference's 0/0 (NaN); a zero numerator gives a signed zero.
template <typename T>
__device__ __forceinline__ T trc_adv(bool neg, int nummask, int den, const T* tab, int j) {
    if (den == 0) return __builtin_nan("");
    if (!nummask) return neg ? T(-0.0) : T(0.0);
    const T v = tab[2 * j + den - 1];
    return neg ? -v : v;
}

template <typename T, int VAR, int NTRC, int TX, int TY, bool DIAG>
__global__ void __launch_bounds__(256)
step_kernel(const Dyn<T> d) {
    constexpr int NT = 256;
    constexpr int EW = TX + 2;
    constexpr int EH = TY + 2;
    constexpr int E = EW * EH;
    constexpr int NO = (TX * TY + NT - 1) / NT;
    constexpr int NTR = NTRC > 0 ? NTRC : 1;
    __shared__ T s_u[E], s_v[E], s_eta[E], s_D[E], s_Pot[E], s_Ed[E], s_MU[E], s_MV[E], s_pll[E], s_plt[E];
    __shared__ T s_CH[NTR][E], s_C[NTR][E];
    __shared__ uint8_t s_m[E];

    const Statics<T>& a = *d.s;
    const int tid = threadIdx.x;
    const int t = xcd_tile(blockIdx.x, d.nblk);
    const int i0 = (t % d.ntx) * TX;
    const int j0 = (t / d.ntx) * TY;
    const int Nx = a.Nx, Ny = a.Ny, P = a.P;
    const bool FN = (VAR == VAR_FN);
    // planes of this step's roles
    T* const st = d.state;
    const long long pl = d.plane;
    const int r0 = d.ring, r1 = d.ring == 2 ? 0 : d.ring + 1, r2 = d.ring == 0 ? 2 : d.ring - 1;
    const T* const su = st + (PL_U + d.cur) * pl;
    const T* const sv = st + (PL_V + d.cur) * pl;
    const T* const se = st + (PL_E + d.cur) * pl;
    const T* const Gu1 = st + (PL_GU + r0) * pl;
    const T* const Gu2 = st + (PL_GU + r1) * pl;
    const T* const Gv1 = st + (PL_GV + r0) * pl;
    const T* const Gv2 = st + (PL_GV + r1) * pl;
    const T* const Ge1 = st + (PL_GE + r0) * pl;
    const T* const Ge2 = st + (PL_GE + r1) * pl;

    // ---- prefetch this thread's AB history (independent of the LDS phases)
    T pg[NO][6];
    T pt[NO][NTR][2];
    if (!DIAG) {
#pragma unroll
        for (int q = 0; q < NO; ++q) {
            const int o = tid + q * NT;
            const int i = i0 + o % TX, j = j0 + o / TX;
            if (o < TX * TY && i < Nx && j < Ny) {
                const size_t off = (size_t)j * P + i;
                pg[q][0] = Gu1[off]; pg[q][1] = Gu2[off];
                pg[q][2] = Gv1[off]; pg[q][3] = Gv2[off];
                pg[q][4] = Ge1[off]; pg[q][5] = Ge2[off];
#pragma unroll
                for (int n = 0; n < NTRC; ++n) {
                    const T* tb = st + (PL_TRC + PL_PER_TRC * n + 2) * pl;
                    pt[q][n][0] = tb[r0 * pl + off];
                    pt[q][n][1] = tb[r1 * pl + off];
                }
            }
        }
    }

    // ---- phase 1: state with a one-cell halo (periodic indices) into LDS
    for (int e = tid; e < E; e += NT) {
        const int lx = e % EW, ly = e / EW;
        const int gi = wrap(i0 + lx - 1, Nx), gj = wrap(j0 + ly - 1, Ny);
        const size_t off = (size_t)gj * P + gi;
        const uint8_t m = a.mask[off];
        const T eta = se[off];
        s_m[e] = m;
        s_u[e] = su[off];
        s_v[e] = sv[off];
        s_eta[e] = eta;
        const T He = fld(a.He, off, gj, OE(m) != 0);
        T D;
        if (FN) D = OE(m) ? (eta + He) : T(1);   // computeD: D = eta + H on ocean eta, 1 elsewhere
        else D = He;                              // linear variants: D aliases H_eta
        s_D[e] = D;
#pragma unroll
        for (int n = 0; n < NTRC; ++n) {
            const T* tb = st + (PL_TRC + PL_PER_TRC * n) * pl;
            const T ch = tb[d.cur * pl + off];
            s_CH[n][e] = ch;
            // C = CH / h (TRC_tracer_advance), except the host-held C right after a hand-over
            s_C[n][e] = ((d.cbuf >> n) & 1) ? tb[5 * pl + off] : ch / D;
        }
    }
    __syncthreads();

    // ---- phase 2: derived fields on the extended tile
    for (int e = tid; e < E; e += NT) {
        const int lx = e % EW, ly = e / EW;
        const int gj = wrap(j0 + ly - 1, Ny);
        const uint8_t m = s_m[e];
        const bool interior = DIAG && lx >= 1 && lx <= TX && ly >= 1 && ly <= TY &&
                              (i0 + lx - 1) < Nx && (j0 + ly - 1) < Ny;
        const size_t doff = interior ? (size_t)(j0 + ly - 1) * P + (i0 + lx - 1) : 0;
        if (lx >= 1 && ly >= 1) {
            const int kw = e - 1, ks = e - EW, ksw = e - EW - 1;
            // Dh = interp4(D, eta2H_noland): points (i,j), (i-1,j-1), (i-1,j), (i,j-1)
            const int m1 = OE(m), m2 = OE(s_m[ksw]), m3 = OE(s_m[kw]), m4 = OE(s_m[ks]);
            const T w = noland_w<T>(m1 + m2 + m3 + m4);
            const T Dh = (((m1 ? w : T(0)) * s_D[e] + (m2 ? w : T(0)) * s_D[ksw]) + (m3 ? w : T(0)) * s_D[kw]) +
                         (m4 ? w : T(0)) * s_D[ks];
            // computeRelVort -> vorticityFromVelocities2D (src/calc_lib.f90:663-678,
            // pder_zonal3D 892-926, pder_meridional3D 965-998)
            T zeta = T(0);
            if (FN) {
                const int gjm = gj == 0 ? Ny - 1 : gj - 1;
                T za_ = T(0), zb = T(0);
                if (OV(s_m[kw]) + OV(m) != 0)
                    za_ = a.za[2 * gj + (OH(m) ? 0 : 1)] * (s_v[e] - s_v[kw]);
                if (OU(m) + OU(s_m[ks]) != 0) {
                    const T bcH = OH(m) ? T(1) : a.lbc;
                    zb = bcH * (a.cu[gj] * s_u[e] - a.cu[gjm] * s_u[ks]) / a.AdT;
                }
                zeta = za_ - zb / a.cH[gj];
            }
            // computePotVort (src/swm_timestep_module.f90:230-257)
            const T pD = Dh > a.eps ? Dh : a.eps;
            const T Pot = FN ? (a.fH[gj] + zeta) / pD : a.fH[gj] / pD;
            s_Pot[e] = Pot;
            // SWM_LateralMixing_compute_p, P_lambda-theta on ocean H points
            T plt = T(0);
            if (a.mix && OH(m)) {
                const T* c = a.pltc + (gj * 4 + OU(m) * 2 + OU(s_m[ks])) * 3;
                plt = Dh * ((c[0] * (s_v[e] - s_v[kw]) + c[1] * s_u[e]) + c[2] * s_u[ks]);
            }
            s_plt[e] = plt;
            if (interior) { a.diag[1][doff] = Dh; a.diag[4][doff] = zeta; a.diag[5][doff] = Pot; }
        }
        if (lx >= 1) {   // computeMassFluxes: MU = Du * u on ocean u
            const T Du = interp_noland2(s_D[e], s_D[e - 1], OE(m), OE(s_m[e - 1]));
            s_MU[e] = OU(m) ? Du * s_u[e] : T(0);
            if (interior) { a.diag[2][doff] = Du; a.diag[7][doff] = s_MU[e]; }
        }
        if (ly >= 1) {   // MV = Dv * v on ocean v
            const T Dv = interp_noland2(s_D[e], s_D[e - EW], OE(m), OE(s_m[e - EW]));
            s_MV[e] = OV(m) ? Dv * s_v[e] : T(0);
            if (interior) { a.diag[3][doff] = Dv; a.diag[8][doff] = s_MV[e]; }
        }
        if (lx <= TX && ly <= TY) {
            const int ke = e + 1, kn = e + EW;
            // computeEdens (src/swm_timestep_module.f90:259-310); land keeps 0
            T Ed = T(0);
            if (OE(m)) {
                if (FN) {
                    const T ue = s_u[ke], uc = s_u[e], vn = s_v[kn], vc = s_v[e];
                    Ed = ((T(0.5) * (ue * ue) + T(0.5) * (uc * uc)) + (T(0.5) * (vn * vn) + T(0.5) * (vc * vc))) / T(2) +
                         a.G * s_eta[e];
                } else {
                    Ed = a.G * s_eta[e];
                }
            }
            s_Ed[e] = Ed;
            // P_lambda-lambda on ocean eta points
            T pll = T(0);
            if (a.mix && OE(m)) {
                const T* c = a.pllc + (gj * 4 + OV(s_m[kn]) * 2 + OV(m)) * 3;
                pll = s_D[e] * ((c[0] * (s_u[ke] - s_u[e]) + c[1] * s_v[kn]) + c[2] * s_v[e]);
            }
            s_pll[e] = pll;
            if (interior) { a.diag[6][doff] = Ed; a.diag[0][doff] = s_D[e]; }
        }
    }
    __syncthreads();

    // ---- phase 3: tendencies, AB integration, clamp, tracers
#pragma unroll
    for (int q = 0; q < NO; ++q) {
        const int o = tid + q * NT;
        const int lx = o % TX + 1, ly = o / TX + 1;
        const int i = i0 + lx - 1, j = j0 + ly - 1;
        if (o >= TX * TY || i >= Nx || j >= Ny) continue;
        const size_t off = (size_t)j * P + i;
        const int k = ly * EW + lx;
        const int ke = k + 1, kw = k - 1, kn = k + EW, ks = k - EW, knw = kn - 1, kse = ks + 1;
        const int jp = j + 1 == Ny ? 0 : j + 1;
        const uint8_t m = s_m[k];
        const int oe = OE(m), ou = OU(m), ov = OV(m), oh = OH(m);
        const T Du = interp_noland2(s_D[k], s_D[kw], oe, OE(s_m[kw]));
        const T Dv = interp_noland2(s_D[k], s_D[ks], oe, OE(s_m[ks]));

        // SWM_forcing_update (src/swm_forcing_module.f90:135-170)
        T Fe = fld(a.Fec, off, j, oe);
        if (a.has_oe) Fe = Fe + (T(2) * (fld(a.oRe, off, j, oe) * d.oc[2] - fld(a.oIe, off, j, oe) * d.os[2])) * T(oe);
        T Fx = fld(a.Fxc, off, j, ou);
        if (a.has_taux && ou) {
            const T tx = fld(a.taux, off, j, true);
            Fx = Fx + (a.has_tau_scale ? a.tau_scale * tx : tx) / (a.RHO0 * Du);
        }
        if (a.has_ox) Fx = Fx + (T(2) * (fld(a.oRx, off, j, ou) * d.oc[0] - fld(a.oIx, off, j, ou) * d.os[0])) * T(ou);
        T Fy = fld(a.Fyc, off, j, ov);
        if (a.has_tauy && ov) {
            const T ty = fld(a.tauy, off, j, true);
            Fy = Fy + (a.has_tau_scale ? a.tau_scale * ty : ty) / (a.RHO0 * Dv);
        }
        if (a.has_oy) Fy = Fy + (T(2) * (fld(a.oRy, off, j, ov) * d.oc[1] - fld(a.oIy, off, j, ov) * d.os[1])) * T(ov);

        // SWM_timestep_nonlinear (src/swm_timestep_module.f90:349-443)
        T Ge = T(0), Gu = T(0), Gv = T(0), Lu = T(0), Lv = T(0);
        if (oe) {
            const T t1 = -(s_MU[ke] - s_MU[k]) / a.dL;
            const T t2 = (s_MV[kn] * a.cv[jp] - s_MV[k] * a.cv[j]) / a.dT;
            Ge = (t1 - t2) / a.ace[j] + Fe;
        }
        if (ou) {
            const T P2 = T(0.5) * s_Pot[kn] + T(0.5) * s_Pot[k];
            const T M4 = ((T(0.25) * s_MV[kn] + T(0.25) * s_MV[kw]) + T(0.25) * s_MV[knw]) + T(0.25) * s_MV[k];
            T g = P2 * M4;
            g = g - ((s_Ed[k] - s_Ed[kw]) / a.edu[j]);
            if (a.quad) {
                const T vi = ((T(0.25) * s_v[kn] + T(0.25) * s_v[kw]) + T(0.25) * s_v[knw]) + T(0.25) * s_v[k];
                const T uc = s_u[k];
                T qf = fld(a.gsu, off, j, true) * sqrt(uc * uc + (vi * vi)) * uc;
                if (a.barotropic) qf = qf / fld(a.Hu, off, j, true);
                g = g - qf;
            }
            if (a.mix) {
                const int ot = (oh + OH(s_m[kn])) < 1 ? 1 : (oh + OH(s_m[kn]));
                const T* c = a.lmu + (j * 2 + ot - 1) * 3;
                Lu = ((c[0] * (s_pll[k] - s_pll[kw]) + c[1] * s_plt[kn]) + c[2] * s_plt[k]) / Du;
                g = g + Lu;
            }
            Gu = g + Fx;
        }
        if (ov) {
            const T P2 = T(0.5) * s_Pot[ke] + T(0.5) * s_Pot[k];
            const T M4 = ((T(0.25) * s_MU[ke] + T(0.25) * s_MU[ks]) + T(0.25) * s_MU[k]) + T(0.25) * s_MU[kse];
            T g = -(P2 * M4);
            g = g - ((s_Ed[k] - s_Ed[ks]) / a.AdT);
            if (a.quad) {
                const T ui = ((T(0.25) * s_u[ke] + T(0.25) * s_u[ks]) + T(0.25) * s_u[k]) + T(0.25) * s_u[kse];
                const T vc = s_v[k];
                T qf = fld(a.gsv, off, j, true) * sqrt(vc * vc + (ui * ui)) * vc;
                if (a.barotropic) qf = qf / fld(a.Hv, off, j, true);
                g = g - qf;
            }
            if (a.mix) {
                const int ot = (oe + OE(s_m[ks])) < 1 ? 1 : (oe + OE(s_m[ks]));
                const T* c = a.lmv + (j * 2 + ot - 1) * 3;
                Lv = ((c[0] * (s_plt[ke] - s_plt[k]) + c[1] * s_pll[k]) + c[2] * s_pll[ks]) / Dv;
                g = g + Lv;
            }
            Gv = g + Fy;
        }

        if (DIAG) {
            a.diag[9][off] = Fx; a.diag[10][off] = Fy; a.diag[11][off] = Fe;
            a.diag[12][off] = Lu; a.diag[13][off] = Lv;
#pragma unroll
            for (int n = 0; n < NTRC; ++n) {
                const T C = s_C[n][k];
                a.tdiag[n][0][off] = C;
                a.tdiag[n][1][off] = s_MU[k] * T(0.5) * (C + s_C[n][kw]);
                a.tdiag[n][2][off] = s_MV[k] * T(0.5) * (C + s_C[n][ks]);
            }
            continue;
        }

        // integrate_AB (src/time_integration_module.f90:74-112); land keeps the
        // previous newest tendency, which the reference never overwrites there.
        const T Gu3 = ou ? Gu : pg[q][1];
        const T Gv3 = ov ? Gv : pg[q][3];
        const T Ge3 = oe ? Ge : pg[q][5];
        T un = s_u[k], vn = s_v[k], en = s_eta[k];
        un = un + d.dtc1 * pg[q][0];
        vn = vn + d.dtc1 * pg[q][2];
        en = en + d.dtc1 * pg[q][4];
        if (d.tstep >= 2) {
            un = un + d.dtc2 * pg[q][1];
            vn = vn + d.dtc2 * pg[q][3];
            en = en + d.dtc2 * pg[q][5];
        }
        if (d.tstep >= 3) {
            un = un + d.dtc3 * Gu3;
            vn = vn + d.dtc3 * Gv3;
            en = en + d.dtc3 * Ge3;
        }
        un = un / fld(a.imu, off, j, ou);
        vn = vn / fld(a.imv, off, j, ov);
        en = en / fld(a.ime, off, j, oe);
        if (FN) {   // SWM_timestep_step: where (eta < minD - H) eta = minD - H
            const T lim = a.minD - fld(a.He, off, j, oe);
            if (en < lim) en = lim;
        }
        st[(PL_U + 1 - d.cur) * pl + off] = un;
        st[(PL_V + 1 - d.cur) * pl + off] = vn;
        st[(PL_E + 1 - d.cur) * pl + off] = en;
        st[(PL_GU + r2) * pl + off] = Gu3;
        st[(PL_GV + r2) * pl + off] = Gv3;
        st[(PL_GE + r2) * pl + off] = Ge3;

        // TRC_tracer_incr + TRC_tracer_integrate (src/tracer_module.f90:154-232)
#pragma unroll
        for (int n = 0; n < NTRC; ++n) {
            T* const tb = st + (PL_TRC + PL_PER_TRC * n) * pl;
            const T CH = s_CH[n][k], CHe = s_CH[n][ke], CHw = s_CH[n][kw], CHn = s_CH[n][kn], CHs = s_CH[n][ks];
            const T C = s_C[n][k], Ce = s_C[n][ke], Cw = s_C[n][kw], Cn = s_C[n][kn], Cs = s_C[n][ks];
            // tracer velocities are the master fields u = 0 + SWM_u (model_advance)
            const T um = T(0) + s_u[k], ume = T(0) + s_u[ke], vm = T(0) + s_v[k], vmn = T(0) + s_v[kn];
            const uint8_t me = s_m[ke], mw = s_m[kw], mn = s_m[kn], ms = s_m[ks];
            const int oee = OE(me), oew = OE(mw), oen = OE(mn), oes = OE(ms);
            const int oue = OU(me), ovn = OV(mn);
            T diff = T(0);
            if (oe) {
                const T kap = a.kappa[n];
                const T hue = interp_noland2(s_D[ke], s_D[k], oee, oe);
                const T hvn = interp_noland2(s_D[kn], s_D[k], oen, oe);
                const T c9 = oue ? a.t9[j] : T(0.0);
                const T c10 = ou ? -a.t9[j] : T(-0.0);
                const T c11 = ovn ? a.t11[j] : T(0.0);
                const T c12 = ov ? -a.t12[j] : T(-0.0);
                diff = ((((kap * hue) * (Ce - C)) * c9 + ((kap * Du) * (C - Cw)) * c10) +
                        ((kap * hvn) * (Cn - C)) * c11) + ((kap * Dv) * (C - Cs)) * c12;
            }
            const T forcing = C * Fe;
            const T c1 = trc_adv<T>(true, oue & oee, oee + oe, a.tA, j);
            const T c2 = trc_adv<T>(false, ou & oew, oew + oe, a.tA, j);
            const T c3 = trc_adv<T>(true, ovn & oen, oen + oe, a.tB, j);
            const T c4 = trc_adv<T>(false, ov & oes, oes + oe, a.tC, j);
            const T c5 = trc_adv<T>(true, oue & oe, oee + oe, a.tA, j);
            const T c6 = trc_adv<T>(false, ou & oe, oew + oe, a.tA, j);
            const T c7 = trc_adv<T>(true, ovn & oe, oen + oe, a.tB, j);
            const T c8 = trc_adv<T>(false, ov & oe, oes + oe, a.tC, j);
            T g = (CHe * ume) * c1 + (CHw * um) * c2;
            g = g + (CHn * vmn) * c3;
            g = g + (CHs * vm) * c4;
            g = g + (CH * ume) * c5;
            g = g + (CH * um) * c6;
            g = g + (CH * vmn) * c7;
            g = g + (CH * vm) * c8;
            g = g + diff;
            g = g - (fld(a.gC[n], off, j, oe) * s_D[k]) * (C - fld(a.C0[n], off, j, oe));
            g = g + forcing;
            T chn = CH + d.dtc1 * pt[q][n][0];
            if (d.tstep >= 2) chn = chn + d.dtc2 * pt[q][n][1];
            if (d.tstep >= 3) chn = chn + d.dtc3 * g;
            chn = chn / fld(a.timpl[n], off, j, oe);
            tb[(1 - d.cur) * pl + off] = chn;
            tb[(2 + r2) * pl + off] = g;
        }
    }
}

#undef OE
#undef OV
#undef OU
#undef OH

// ---------------------------------------------------------------------------
// Red-black SOR (src/ElSolv_SOR.f90:150-222)
// ---------------------------------------------------------------------------
template <typename T>
struct SorArgs {
    int Nx, Ny, P, color, sweep;
    T* chi;
    const T* B;
    const uint8_t* mask;
    const double* relax;    // omega per half sweep (reference mode), indexed 2*(sweep-1)+color, clamped
    int nrelax;
    double omega;           // fixed-omega mode when relax == nullptr
    const T *c1t, *c3t, *c4t;   // [Ny] 1/x, cv(j+1)/y, cv(j)/y
    const T* c5t;               // [Ny][3][2][2] c5 by (ou(i+1)+ou(i), ov(j+1), ov(j))
    double* partial;            // per-block partial sums of res^2 (or r^2)
    const int* done;            // set once converged: later sweeps return at once
    int residual_mode;          // 0: sum res^2 (reference), 1: sum (L chi - B)^2
};

// Each thread updates one cell of the colour; row j holds the colour's cells
// at i = (j + colour + 1) % 2 + 2k (colour 0 = (i+j) even in 1-based indices).
template <typename T>
__global__ void __launch_bounds__(256) sor_sweep(const SorArgs<T> a) {
    __shared__ double red[256 / 64];
    if (*a.done) return;
    const int half = (a.Nx + 1) / 2;
    const long idx = (long)blockIdx.x * 256 + threadIdx.x;
    const int j = (int)(idx / half);
    const int kk = (int)(idx % half);
    double acc = 0.0;
    if (j < a.Ny) {
        // 1-based (i+j) even <=> 0-based (i+j) even
        const int i = 2 * kk + ((j + a.color) & 1);
        if (i < a.Nx) {
            const size_t off = (size_t)j * a.P + i;
            const uint8_t m = a.mask[off];
            if (m & OC_E) {
                const int ip = i + 1 == a.Nx ? 0 : i + 1, im = i == 0 ? a.Nx - 1 : i - 1;
                const int jp = j + 1 == a.Ny ? 0 : j + 1, jm = j == 0 ? a.Ny - 1 : j - 1;
                const uint8_t mip = a.mask[(size_t)j * a.P + ip], mjp = a.mask[(size_t)jp * a.P + i];
                const int ouip = (mip >> 1) & 1, ou = (m >> 1) & 1, ovjp = (mjp >> 2) & 1, ov = (m >> 2) & 1;
                const T c1 = ouip ? a.c1t[j] : T(0);
                const T c2 = ou ? a.c1t[j] : T(0);
                const T c3 = ovjp ? a.c3t[j] : T(0);
                const T c4 = ov ? a.c4t[j] : T(0);
                const T c5 = a.c5t[((j * 3 + ouip + ou) * 2 + ovjp) * 2 + ov];
                T* chi = a.chi;
                const T x = chi[off];
                const T lhs = (((c1 * chi[(size_t)j * a.P + ip] + c2 * chi[(size_t)j * a.P + im]) +
                                c3 * chi[(size_t)jp * a.P + i]) + c4 * chi[(size_t)jm * a.P + i]) + c5 * x;
                const T r = lhs - a.B[off];
                T res;
                if (a.relax) {
                    int h = 2 * (a.sweep - 1) + a.color;
                    if (h >= a.nrelax) h = a.nrelax - 1;
                    res = T(a.relax[h]) * r / c5;
                } else {
                    res = T(a.omega) * r / c5;
                }
                chi[off] = x - res;
                const double q = a.residual_mode ? (double)r : (double)res;
                acc = q * q;
            }
        }
    }
    // fixed-order block reduction: wavefront butterfly then the 4 wave sums in order
    for (int s = 32; s >= 1; s >>= 1) acc += __shfl_xor(acc, s, 64);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0) red[wv] = acc;
    __syncthreads();
    if (threadIdx.x == 0) a.partial[blockIdx.x] = ((red[0] + red[1]) + red[2]) + red[3];
}

// Sum the block partials in a fixed order (one block), test convergence.
// mode 0: sqrt(sum) < eps (reference); mode 1: sqrt(sum)/bnorm < eps.
__global__ void __launch_bounds__(256)
sor_check(const double* partial, int nblk, int sweep, double eps, double bnorm, int mode,
          int* done, long long* iters, double* norm_out) {
    __shared__ double red[4];
    if (*done) return;
    double acc = 0.0;
    for (int b = threadIdx.x; b < nblk; b += 256) acc += partial[b];
    for (int s = 32; s >= 1; s >>= 1) acc += __shfl_xor(acc, s, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        const double tot = ((red[0] + red[1]) + red[2]) + red[3];
        double nrm = sqrt(tot);
        if (mode == 1) nrm = bnorm > 0.0 ? nrm / bnorm : nrm;
        *norm_out = nrm;
        *iters = sweep;
        if (nrm < eps) *done = 1;
    }
}

// computeDivergence(u, v, div) (src/calc_lib.f90:607-622)
template <typename T>
__global__ void __launch_bounds__(256)
divergence_kernel(int Nx, int Ny, int P, const T* u, const T* v, const uint8_t* mask, const T* dza,
                  const T* cv, const T* ce, T AdT, T lbc, T* div) {
    const long idx = (long)blockIdx.x * 256 + threadIdx.x;
    if (idx >= (long)Nx * Ny) return;
    const int i = (int)(idx % Nx), j = (int)(idx / Nx);
    const int ip = i + 1 == Nx ? 0 : i + 1, jp = j + 1 == Ny ? 0 : j + 1;
    const size_t off = (size_t)j * P + i;
    const uint8_t m = mask[off], mip = mask[(size_t)j * P + ip], mjp = mask[(size_t)jp * P + i];
    T za = T(0), zb = T(0);
    if (((m >> 1) & 1) + ((mip >> 1) & 1) != 0)
        za = dza[2 * j + ((m & OC_E) ? 0 : 1)] * (u[(size_t)j * P + ip] - u[off]);
    if (((mjp >> 2) & 1) + ((m >> 2) & 1) != 0) {
        const T bc = (m & OC_E) ? T(1) : lbc;
        zb = bc * (cv[jp] * v[(size_t)jp * P + i] - cv[j] * v[off]) / AdT;
    }
    div[off] = za + zb / ce[j];
}

}  // namespace swm
