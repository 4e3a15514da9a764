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
enum { VAR_FN = 1, VAR_MS = 2, VAR_SOR = 3 };

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

// Packed per-row values that phase 3 of the tile kernels reads, RT_N per row
// (Statics::rowtab): built on the device from the per-row tables and the
// row-compressed fields whenever the statics are written.  A field stored as
// a full array leaves its two entries 0 and is read from the array.  RT_R*
// are the correctly rounded reciprocals of the divisors (dvr).  Tracer n has
// RT_PER_TRC entries from RT_TRC + n * RT_PER_TRC: timpl, gC, C0, 1 / timpl
// (ocean, land each).
enum : int {
    RT_CV = 0, RT_ACE = 1, RT_EDU = 2, RT_LMU = 3, RT_LMV = 9, RT_T9 = 15, RT_T11 = 16, RT_T12 = 17,
    RT_TA = 18, RT_TB = 20, RT_TC = 22, RT_FEC = 24, RT_FXC = 26, RT_FYC = 28, RT_IMU = 30, RT_IMV = 32,
    RT_IME = 34, RT_HE = 36, RT_TAUX = 38, RT_TAUY = 40, RT_RACE = 42, RT_REDU = 43, RT_RIMU = 44,
    RT_RIMV = 46, RT_RIME = 48, RT_TRC = 50, RT_PER_TRC = 8, RT_N = RT_TRC + 2 * RT_PER_TRC
};

// Static inputs of the step kernel, resident in device memory and read through
// one pointer (scalar loads on use), so the per-launch argument block stays
// small and the kernel does not spill uniform values out of SGPRs.
template <typename T>
struct Statics {
    int Nx, Ny, P;
    int mix, quad, barotropic, has_tau_scale;
    int carry;           // FP32 state carries its bfloat16 low part (lo planes read and written)
    int has_taux, has_tauy, has_ox, has_oy, has_oe;
    T G, A, dL, dT, RHO0, minD, lbc, tau_scale, AdT, eps;
    T rdL, rdT, rAdT;    // 1/dL, 1/dT, 1/AdT (correctly rounded, for dvr)
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
    const T* rowtab;         // [Ny][RT_N] packed per-row values of phase 3
    // tracers
    T kappa[2];
    Fld<T> C0[2], gC[2], timpl[2];
    // LINEARISED_MEAN_STATE basic state u_bs, v_bs (u, v grids) and zeta_bs (H grid)
    Fld<T> ubs, vbs, zbs;
    // time-interpolated forcing streams: kind 0 = windstress, 1 = custom or heating
    // (added as is); component 0 = x (ocean u), 1 = y (ocean v), 2 = eta
    int nstream, sdiv;   // sdiv: windstress divides by RHO0 * D_face (BAROTROPIC or FULLY_NONLINEAR)
    int skind[4], scomp[4];
    const T* srec[4];    // records of each stream, one plane apart
    // DYNFROMFILE datasets of u, v (records, one plane apart; null when absent)
    const T *dyn_u, *dyn_v;
    // diagnostic outputs (materialisation kernel only)
    T* diag[14];      // D Dh Du Dv zeta Pot EDens MU MV F_x F_y F_eta latmix_u latmix_v
    T* tdiag[2][3];   // C uhc vhc per tracer
};

// Fill Statics::rowtab ([Ny][RT_N]) from the tables and the row-compressed
// fields; entries of absent tracers and of full-array fields are 0.
template <typename T>
__global__ void __launch_bounds__(256) pack_rowtab(const Statics<T>* __restrict__ sp, T* __restrict__ out, int Ny) {
    const int x = blockIdx.x * 256 + threadIdx.x;
    if (x >= Ny * RT_N) return;
    const Statics<T>& a = *sp;
    const int j = x / RT_N, k = x % RT_N;
    auto rows = [&](const Fld<T>& f, int s) -> T { return f.rows ? f.rows[2 * j + s] : T(0); };
    auto rcp = [](T b) -> T { return b != T(0) ? T(1) / b : T(0); };   // correctly rounded
    T v;
    if (k == RT_CV) v = a.cv[j];
    else if (k == RT_ACE) v = a.ace[j];
    else if (k == RT_EDU) v = a.edu[j];
    else if (k < RT_LMV) v = a.lmu[6 * j + k - RT_LMU];
    else if (k < RT_T9) v = a.lmv[6 * j + k - RT_LMV];
    else if (k == RT_T9) v = a.t9[j];
    else if (k == RT_T11) v = a.t11[j];
    else if (k == RT_T12) v = a.t12[j];
    else if (k < RT_TB) v = a.tA[2 * j + k - RT_TA];
    else if (k < RT_TC) v = a.tB[2 * j + k - RT_TB];
    else if (k < RT_FEC) v = a.tC[2 * j + k - RT_TC];
    else if (k < RT_FXC) v = rows(a.Fec, k - RT_FEC);
    else if (k < RT_FYC) v = rows(a.Fxc, k - RT_FXC);
    else if (k < RT_IMU) v = rows(a.Fyc, k - RT_FYC);
    else if (k < RT_IMV) v = rows(a.imu, k - RT_IMU);
    else if (k < RT_IME) v = rows(a.imv, k - RT_IMV);
    else if (k < RT_HE) v = rows(a.ime, k - RT_IME);
    else if (k < RT_TAUX) v = rows(a.He, k - RT_HE);
    else if (k < RT_TAUY) v = rows(a.taux, k - RT_TAUX);
    else if (k < RT_RACE) v = rows(a.tauy, k - RT_TAUY);
    else if (k == RT_RACE) v = rcp(a.ace[j]);
    else if (k == RT_REDU) v = rcp(a.edu[j]);
    else if (k < RT_RIMV) v = rcp(rows(a.imu, k - RT_RIMU));
    else if (k < RT_RIME) v = rcp(rows(a.imv, k - RT_RIMV));
    else if (k < RT_TRC) v = rcp(rows(a.ime, k - RT_RIME));
    else {
        const int n = (k - RT_TRC) / RT_PER_TRC, r = (k - RT_TRC) % RT_PER_TRC;
        v = rows(r < 2 || r >= 6 ? a.timpl[n] : (r < 4 ? a.gC[n] : a.C0[n]), r & 1);
        if (r >= 6) v = rcp(v);
    }
    out[x] = v;
}

// State planes in one allocation: u, v, eta at two levels, three AB slots per
// field, the low parts of u, v, eta (FP32 state), then per tracer CH (two
// levels), G_CH (three slots), the C buffer and the low part of CH.
enum : int {
    PL_U = 0, PL_V = 2, PL_E = 4, PL_GU = 6, PL_GV = 9, PL_GE = 12, PL_LO = 15, PL_TRC = 18, PL_PER_TRC = 7,
    PL_TLO = 6   // tracer plane of the low part of CH
};

// FP32 state with a carried low part.  A cell's value is hi + lo: hi in the
// float plane the step reads, lo rounded to bfloat16 in a plane of its own
// (|lo| <= ulp(hi)/2, so the pair holds about 32 significant bits).  The
// step adds its increment to the pair with an exact two-sum and renormalises,
// so the rounding of hi is carried to the next step instead of lost; without
// it the rounding of the many small AB increments into hi accumulates to
// O(steps * 2^-24) relative error in long runs.  The stored lo loses at most
// 2^-33 |hi| per step.
__device__ __forceinline__ float lo_dec(unsigned short h) { return __uint_as_float((unsigned)h << 16); }
__device__ __forceinline__ unsigned short lo_enc(float x) {   // round to nearest even
    const unsigned b = __float_as_uint(x);
    return (unsigned short)((b + 0x7fffu + ((b >> 16) & 1u)) >> 16);
}
struct Sum2f {
    float s, e;
};
__device__ __forceinline__ Sum2f two_sum(float a, float b) {   // s + e == a + b exactly
    const float s = a + b;
    const float bp = s - a;
    return {s, (a - (s - bp)) + (b - bp)};
}
// (hi + lo_in + inc) / impl with kap = 1 - 1/impl (0 where there is no
// damping): the new (hi, lo).  The FP32 statics hold kap, not impl: impl =
// 1 + dt*gamma is 1 + O(1e-5) for a linear drag, and a float keeps only a
// few bits of its distance from 1 (ulp(1) = 1.2e-7), which changes the drag
// by up to 1 % and, through the flow it damps, the amplitude of every field
// by a relative O(1e-8) per step.  x * (1 - kap) is formed as x - x*kap with
// the product's FMA error and the difference's two-sum error carried in lo.
__device__ __forceinline__ float carry_step(float hi, unsigned short lo_in, float inc, float kap,
                                            unsigned short& lo_out) {
    const Sum2f t = two_sum(hi, inc + lo_dec(lo_in));
    float q = t.s, e = t.e;
    if (kap != 0.0f) {
        const float p = t.s * kap;
        const float ep = __builtin_fmaf(t.s, kap, -p);   // t.s * kap == p + ep
        const Sum2f w = two_sum(t.s, -p);
        q = w.s;
        e = (w.e - ep) + (e - e * kap);
    }
    const float h = q + e;
    const bool fin = __builtin_isfinite(q);   // Inf / NaN: the plain float result, no low part
    lo_out = fin ? lo_enc(e - (h - q)) : (unsigned short)0;
    return fin ? h : q;
}
template <typename T>
__device__ __forceinline__ const unsigned short* lo_plane(const T* st, long long pl, int k) {
    return (const unsigned short*)(st + (long long)k * pl);
}

// One forcing stream's place on its time axis for a step: records o0 and o1
// (element offsets) bracket the model time; the value is rec[o0] when exact,
// else rec[o0] + a*(rec[o1] - rec[o0])/b with a = t - t0, b = t1 - t0.
template <typename T>
struct SStep {
    long long o0, o1;
    int exact;
    T a, b;
};

// Per-launch arguments (by value).
template <typename T>
struct Dyn {
    const Statics<T>* s;
    T* state;
    long long plane;     // elements per plane (Ny * P)
    int ntx, nblk;
    int ty;              // rows per wavefront (marching kernel)
    int tpb;             // tiles per block (persistent tile kernel)
    int cur, ring;       // ping-pong level of N0; ring slot of G(1)
    int tstep;           // min(3, itt)
    int cbuf;            // bit n: tracer n reads its C buffer (first step after hand-over)
    T dtc1, dtc2, dtc3;  // dt*ab_coeff(ti, tstep)
    T oc[3], os[3];      // cos/sin(omega*itt*dt) of the oscillating streams (x, y, eta)
    SStep<T> ss[4];      // time-interpolated forcing streams
    SStep<T> ds[2];      // DYNFROMFILE u, v datasets at the master fields' time
};

// A DYNFROMFILE dataset at element off: getChunkData interpolated in time as
// the forcing streams are (src/memchunk_module.f90:214-243).
template <typename T>
__device__ __forceinline__ T dyn_val(const T* r, const SStep<T>& st, size_t off) {
    const T y0 = ((const __attribute__((address_space(1))) T*)(r + st.o0))[off];
    if (st.exact) return y0;
    const T y1 = ((const __attribute__((address_space(1))) T*)(r + st.o1))[off];
    return y0 + (st.a * (y1 - y0)) / st.b;
}

// weight = 1 / max(sum of masks, 1) of the *_noland interpolators
// (src/calc_lib.f90:259-262, 326-332): the exact values of those divisions.
template <typename T>
__device__ __forceinline__ T noland_w(int s) {
    return s <= 1 ? T(1) : (s == 2 ? T(0.5) : (s == 3 ? T(1) / T(3) : T(0.25)));
}

// x mod n for x in [-n, 2n)
__device__ __forceinline__ int wrap1(int x, int n) { return x < 0 ? x + n : (x >= n ? x - n : x); }

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

// a * b + c.  FP64: the product and the sum rounded separately, as written
// (the file is compiled without contraction); IEEE addition is commutative, so
// a sum of products written as nested fm's, p1 + p2 + ... = fm(.., fm(b2, c2,
// p1)), keeps the bits of the sum in the reference's order.  FP32: one fused
// multiply-add (FP32 storage carries an error bar, not bitwise parity).
__device__ __forceinline__ double fm(double a, double b, double c) { return a * b + c; }
__device__ __forceinline__ float fm(float a, float b, float c) { return __builtin_fmaf(a, b, c); }

// interp2 with *_noland weights of D between two eta points (eta2u / eta2v)
template <typename T>
__device__ __forceinline__ T interp_noland2(T D1, T D2, int m1, int m2) {
    const T w = noland_w<T>(m1 + m2);
    return fm(m2 ? w : T(0), D2, (m1 ? w : T(0)) * D1);
}

// Arithmetic of the step kernels.  FP64 evaluates every division and
// multiply-add exactly as written (the file is compiled without contraction),
// which keeps FP64 bit-identical to the reference order.  FP32 storage has a
// stated error bar instead (1e-4 relative max-norm on eta against FP64), so it
// divides through the hardware reciprocal and fuses multiply-adds.
__device__ __forceinline__ double dv(double a, double b) { return a / b; }
__device__ __forceinline__ float dv(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
// a / b where y = 1 / b correctly rounded is at hand (constant and per-row
// divisors).  FP64 gives exactly a / b: q0 = a*y, a remainder step
// r = a - b*q (exact with an FMA) and q + r*y makes q faithful, and a second
// step rounds it correctly (Markstein's theorem, P. W. Markstein, IBM J. Res.
// Dev. 34(1):111-119, 1990); zero and infinite quotients are q0, which
// carries their sign.  Five FMA-pipe operations and a class test in place of
// the scaled reciprocal iteration of a full division.
__device__ __forceinline__ double dvr(double a, double b, double y) {
    const double q0 = a * y;
    double r = __builtin_fma(-b, q0, a);
    const double q1 = __builtin_fma(r, y, q0);
    r = __builtin_fma(-b, q1, a);
    const double q2 = __builtin_fma(r, y, q1);
    return __builtin_amdgcn_class(q0, 0x264) ? q0 : q2;   // 0x264: -inf, -0, +0, +inf
}
__device__ __forceinline__ float dvr(float a, float b, float y) { return a * y; }
__device__ __forceinline__ double madd(double acc, double x, double y) { return acc + x * y; }
__device__ __forceinline__ float madd(float acc, float x, float y) { return __builtin_fmaf(x, y, acc); }

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

// Loads of the step kernels: per-lane values through a global pointer with
// a 32-bit lane index (base address stays in SGPRs), and uniform values of the
// read-only per-row tables through the constant address space (scalar loads).
template <typename T>
__device__ __forceinline__ T ldg(const T* p, unsigned i) {
    return ((const __attribute__((address_space(1))) T*)p)[i];
}
template <typename T>
__device__ __forceinline__ void stg(T* p, unsigned i, T v) {
    ((__attribute__((address_space(1))) T*)p)[i] = v;
}
template <typename T>
__device__ __forceinline__ T ldc(const T* p, size_t i) {
    return ((const __attribute__((address_space(4))) T*)p)[i];
}
// fld() for a uniform row: ro = row offset, j = row index
template <typename T>
__device__ __forceinline__ T fld_r(const Fld<T>& f, size_t ro, unsigned gi, int j, bool ocean) {
    if (f.full) return ldg(f.full + ro, gi);
    const T o = ldc(f.rows, 2 * (size_t)j), l = ldc(f.rows, 2 * (size_t)j + 1);
    return ocean ? o : l;
}
// trc_adv() for a uniform row
template <typename T>
__device__ __forceinline__ T trc_adv_r(bool neg, int nummask, int den, const T* tab, int j) {
    const T v1 = ldc(tab, 2 * (size_t)j), v2 = ldc(tab, 2 * (size_t)j + 1);
    if (den == 0) return __builtin_nan("");
    if (!nummask) return neg ? T(-0.0) : T(0.0);
    const T v = den == 1 ? v1 : v2;
    return neg ? -v : v;
}

// LDS of one tile: the state with a one-cell halo, the derived fields and the
// per-row coefficient tables of its rows.  eta shares its array with Pot (each
// cell reads its own eta before writing its Pot), and the mass fluxes MU, MV
// are formed from Du, Dv and u, v where they are needed, which keeps the FP64
// tile under 32 KiB (five blocks per CU).  NTR is the largest tracer count the
// kernel handles.
template <typename T, int NTR, int E, int EH, int ZE = 1, int DUE = 1, int DVE = 1, int RTR = 1, int RTK_ = 1>
struct TileLds {
    static constexpr int RTK = RTK_;
    T u[E], v[E], D[E], Pot[E], Ed[E], pll[E], plt[E];   // Pot holds eta until phase 2 replaces it
    T zt[ZE], et[ZE];    // LINEARISED_MEAN_STATE (ZE = E): relative vorticity and eta for phase 3
    T Du[DUE], Dv[DVE];  // interp2(D, eta2u/eta2v_noland) of every halo cell (when DUE / DVE = E)
    T CH[NTR][E], C[NTR][E];
    // per-row tables of the EH tile rows (cu also of the row above them)
    T za[EH][2], cu[EH + 1], cH[EH], rcH[EH], fH[EH], pltc[EH][12], pllc[EH][12];
    T rt[RTR][RTK_];     // rowtab rows of the output rows and the row after them (RowTabC)
    int ocean[4];        // per wavefront: all of its halo cells are ocean at every point (tile_body)
    uint8_t m[E];
};

// Phase 3's per-row values: tiles whose wavefronts span several rows stage
// the rowtab entries of their output rows and of the row after them in LDS
// (one coalesced copy per tile instead of a vector load per value and lane);
// 64-wide tiles with one output per thread (a row per wavefront) read them
// with scalar loads.  K = entries a kernel with NTRC tracers reads.
template <int TX, int TY, int NTRC>
struct RowTabC {
    static constexpr int NO = (TX * TY + 255) / 256;
    static constexpr bool UROW = TX == 64 && NO == 1;
    static constexpr int K = RT_TRC + RT_PER_TRC * NTRC;
    static constexpr int R = TY + 1;
    static constexpr int LR = UROW ? 1 : R, LK = UROW ? 1 : K;
    static constexpr int NQ = UROW ? 1 : (R * K + 255) / 256;
};

// Tiles that keep Du and Dv in LDS, computed once per halo cell in phase 2
// instead of at each stencil point in phase 3: the FP32 tiles, which have
// LDS to spare at their occupancy (measured 0.8 % faster at 1024^2).  FP64
// recomputes them (no faster with Du kept, and the 64-wide tiles would drop
// to four blocks per CU).
template <typename T, int TX, int TY>
struct LdsDuDv {
    static constexpr int E = (TX + 2) * (TY + 2);
    static constexpr bool SDU = sizeof(T) == 4;
    static constexpr bool SDV = sizeof(T) == 4;
    static constexpr int DUE = SDU ? E : 1, DVE = SDV ? E : 1;
};

// Global loads and stores through an address-space-1 pointer, and the statics
// through the constant address space.  Pointers that reach a kernel through
// memory (the batched kernel reads them from its member table) are otherwise
// generic and every access would become a flat instruction.
template <typename T>
__device__ __forceinline__ T gld(const T* p, size_t i) {
    return ((const __attribute__((address_space(1))) T*)p)[i];
}
template <typename T>
__device__ __forceinline__ void gst(T* p, size_t i, T v) {
    ((__attribute__((address_space(1))) T*)p)[i] = v;
}
template <typename T>
using CStatics = const __attribute__((address_space(4))) Statics<T>;

// A static field at element offset `at` of row `row` (full array, or the
// row's ocean / land value), read through the global address space.
template <typename F>
__device__ __forceinline__ auto fldg(const F& f, size_t at, int row, bool ocean) {
    return f.full ? gld(f.full, at) : gld(f.rows, 2 * (size_t)row + (ocean ? 0 : 1));
}

// Time-interpolated forcing streams, after the descriptor's windstress and
// oscillating streams: SWM_forcing_processWindstress / processCustomForcing /
// processHeating (src/swm_forcing_module.f90:274-324, 395-434, 535-552) on
// getChunkData (src/memchunk_module.f90:214-243), which interpolates between
// the bracketing records as interp1D_array2D, y0 + (t-t0)*(y1-y0)/(t1-t0)
// (src/calc_lib.f90:386-397).  off is the cell's offset within a plane.
template <typename T, typename S>
__device__ __forceinline__ void apply_streams(S& a, const Dyn<T>& d, size_t off, int ou, int ov, int oe, T Du, T Dv,
                                              T& Fx, T& Fy, T& Fe) {
    const int ns = a.nstream;
    // unrolled over the fixed maximum: every d.ss[n] index is a constant, so a
    // Dyn built in registers (the batched kernel) is not demoted to scratch
#pragma unroll
    for (int n = 0; n < 4; ++n) {
        if (n >= ns) break;
        const int comp = a.scomp[n];
        if (!(comp == 0 ? ou : comp == 1 ? ov : oe)) continue;
        const T* r = a.srec[n];
        const T y0 = gld(r + d.ss[n].o0, off);
        T y = y0;
        if (!d.ss[n].exact) y = y0 + dv(d.ss[n].a * (gld(r + d.ss[n].o1, off) - y0), d.ss[n].b);
        if (a.skind[n] == 0) {
            const T num = a.has_tau_scale ? a.tau_scale * y : y;
            y = a.sdiv ? dv(num, a.RHO0 * (comp == 0 ? Du : Dv)) : dv(num, a.RHO0);
        }
        if (comp == 0) Fx = Fx + y;
        else if (comp == 1) Fy = Fy + y;
        else Fe = Fe + y;
    }
}

// Registers of one tile between its loads and their LDS stores: the state with
// halo (NE cells per thread), the per-row tables and the AB history of the
// thread's NO output cells.
template <typename T, int NE, int NO, int NTR, int NQ = 1>
struct TileRegs {
    T rt[NQ];    // rowtab values for the LDS row table (RowTabC)
    int m[NE];
    int ocean;   // every halo cell of this thread is ocean at all four points
    T u[NE], v[NE], e[NE], Ho[NE], Hl[NE], ch[NTR][NE], cb[NTR][NE];
    T tz0, tz1, tcu, tcH, tfH, tcu0, tpc;
    T g[NO][6];
    T gt[NO][NTR][2];
    T eo[NO];    // eta at the outputs
    unsigned short lo[NO][3 + NTR];   // low parts of u, v, eta, CH at the outputs (FP32 state)
};

template <int TX, int TY>
struct TileShapeC {
    static constexpr int NT = 256, EW = TX + 2, EH = TY + 2, E = EW * EH;
    static constexpr int NO = (TX * TY + NT - 1) / NT, NE = (E + NT - 1) / NT;
};

// Issue the loads of the outputs' AB history (2 slots per field and tracer)
// and eta at the outputs into R.
template <typename T, int VAR, int NTRC, int TX, int TY, bool DIAG, typename RG>
__device__ __forceinline__ void tile_load_hist(const Dyn<T>& d, const int t, RG& R) {
    if (DIAG) return;
    constexpr int NT = 256;
    constexpr int NO = (TX * TY + NT - 1) / NT;
    CStatics<T>& a = *(CStatics<T>*)d.s;
    const int tid = threadIdx.x;
    const int i0 = (t % d.ntx) * TX;
    const int j0 = (t / d.ntx) * TY;
    const int Nx = a.Nx, Ny = a.Ny, P = a.P;
    T* const st = d.state;
    const long long pl = d.plane;
    const int r0 = d.ring, r1 = d.ring == 2 ? 0 : d.ring + 1;
    const T* const se = st + (PL_E + d.cur) * pl;
    const T* const Gu1 = st + (PL_GU + r0) * pl;
    const T* const Gu2 = st + (PL_GU + r1) * pl;
    const T* const Gv1 = st + (PL_GV + r0) * pl;
    const T* const Gv2 = st + (PL_GV + r1) * pl;
    const T* const Ge1 = st + (PL_GE + r0) * pl;
    const T* const Ge2 = st + (PL_GE + r1) * pl;
    auto& pg = R.g;
    auto& pt = R.gt;
    {

#pragma unroll
        for (int q = 0; q < NO; ++q) {
            const int o = tid + q * NT;
            const int i = i0 + o % TX, j = j0 + o / TX;
            if (o < TX * TY && i < Nx && j < Ny) {
                const size_t off = (size_t)j * P + i;
                pg[q][0] = gld(Gu1, off); pg[q][1] = gld(Gu2, off);
                pg[q][2] = gld(Gv1, off); pg[q][3] = gld(Gv2, off);
                pg[q][4] = gld(Ge1, off); pg[q][5] = gld(Ge2, off);
                R.eo[q] = gld(se, off);
#pragma unroll
                for (int n = 0; n < NTRC; ++n) {
                    const T* tb = st + (PL_TRC + PL_PER_TRC * n + 2) * pl;
                    pt[q][n][0] = gld(tb, r0 * pl + off);
                    pt[q][n][1] = gld(tb, r1 * pl + off);
                }
                if constexpr (sizeof(T) == 4) {
                    if (a.carry) {
#pragma unroll
                        for (int f = 0; f < 3; ++f) R.lo[q][f] = gld(lo_plane(st, pl, PL_LO + f), off);
#pragma unroll
                        for (int n = 0; n < NTRC; ++n)
                            R.lo[q][3 + n] = gld(lo_plane(st, pl, PL_TRC + PL_PER_TRC * n + PL_TLO), off);
                    } else {
#pragma unroll
                        for (int f = 0; f < 3 + NTRC; ++f) R.lo[q][f] = 0;
                    }
                }
            }
        }
    }

}

// Issue every global load of tile t: the state with its halo, the per-row
// tables and the AB history of the outputs.
template <typename T, int VAR, int NTRC, int TX, int TY, bool DIAG, bool HIST = true, typename RG>
__device__ __forceinline__ void tile_load(const Dyn<T>& d, const int t, RG& R) {
    constexpr int NT = 256;
    constexpr int EW = TX + 2;
    constexpr int EH = TY + 2;
    constexpr int E = EW * EH;
    constexpr int NO = (TX * TY + NT - 1) / NT;
    constexpr int NTR = NTRC > 0 ? NTRC : 1;
    CStatics<T>& a = *(CStatics<T>*)d.s;
    const int tid = threadIdx.x;
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

    if (HIST) tile_load_hist<T, VAR, NTRC, TX, TY, DIAG>(d, t, R);

    // tiles no larger than the grid: every halo and table index is within one
    // period of it
    const bool one = Nx >= TX && Ny >= TY;
    // ---- phase 1: state with a one-cell halo (periodic indices) into LDS.
    // Every global load of the block is issued before the first LDS store, so
    // one memory latency covers the whole tile and its per-row tables.
    constexpr int NE = (E + NT - 1) / NT;
    auto& rm = R.m;
    auto& ru = R.u;
    auto& rv = R.v;
    auto& re = R.e;
    auto& rHo = R.Ho;
    auto& rHl = R.Hl;
    auto& rch = R.ch;
    auto& rcb = R.cb;
    R.ocean = 1;
#pragma unroll
    for (int q = 0; q < NE; ++q) {
        const int e = tid + q * NT;
        if (e < E) {
            const int lx = e % EW, ly = e / EW;
            int gi = i0 + lx - 1, gj = j0 + ly - 1;
            if (one) {   // uniform: the halo is at most one period outside the grid
                gi = wrap1(gi, Nx);
                gj = wrap1(gj, Ny);
            } else {
                gi = wrap(gi, Nx);
                gj = wrap(gj, Ny);
            }
            const size_t off = (size_t)gj * P + gi;
            rm[q] = gld(a.mask, off);
            R.ocean &= rm[q] == 15;
            ru[q] = gld(su, off);
            rv[q] = gld(sv, off);
            re[q] = gld(se, off);
            // H_eta at an ocean (rHo) and a land (rHl) point of this cell's row
            rHo[q] = a.He.full ? gld(a.He.full, off) : gld(a.He.rows, 2 * (size_t)gj);
            rHl[q] = a.He.full ? rHo[q] : gld(a.He.rows, 2 * (size_t)gj + 1);
#pragma unroll
            for (int n = 0; n < NTRC; ++n) {
                const T* tb = st + (PL_TRC + PL_PER_TRC * n) * pl;
                rch[n][q] = gld(tb, d.cur * pl + off);
                rcb[n][q] = ((d.cbuf >> n) & 1) ? gld(tb, 5 * pl + off) : T(0);
            }
        }
    }
    // per-row tables of rows j0-1 .. j0+TY for phase 2 (cu also of row j0-2)
    T& tz0 = R.tz0;
    T& tz1 = R.tz1;
    T& tcu = R.tcu;
    T& tcH = R.tcH;
    T& tfH = R.tfH;
    T& tcu0 = R.tcu0;
    T& tpc = R.tpc;
    if (tid < EH) {
        const int gj = wrap(j0 + tid - 1, Ny);
        tz0 = gld(a.za, 2 * gj);
        tz1 = gld(a.za, 2 * gj + 1);
        tcu = gld(a.cu, gj);
        tcH = gld(a.cH, gj);
        tfH = gld(a.fH, gj);
    }
    if (tid == NT - 1) tcu0 = gld(a.cu, wrap(j0 - 2, Ny));
    const bool has_pc = a.mix && tid < EH * 24;
    if (has_pc) {
        const int r = tid / 24, c = tid % 24;
        const int gj = wrap(j0 + r - 1, Ny);
        tpc = c < 12 ? gld(a.pltc, gj * 12 + c) : gld(a.pllc, gj * 12 + c - 12);
    }
    // rowtab rows j0 .. j0+TY for phase 3 (tiles that keep them in LDS)
    using RTC = RowTabC<TX, TY, NTRC>;
    if constexpr (!RTC::UROW) {
#pragma unroll
        for (int q = 0; q < RTC::NQ; ++q) {
            const int x = tid + q * NT;
            if (x < RTC::R * RTC::K)
                R.rt[q] = gld(a.rowtab, (size_t)(one ? wrap1(j0 + x / RTC::K, Ny) : wrap(j0 + x / RTC::K, Ny)) * RT_N +
                                            x % RTC::K);
        }
    }
}

// Derive D and C from the loaded state and write the tile to LDS.
template <typename T, int VAR, int NTRC, int TX, int TY, bool DIAG, typename RG, typename L>
__device__ __forceinline__ void tile_stage(const Dyn<T>& d, const int t, const RG& R, L& lds) {
    constexpr int NT = 256;
    constexpr int EW = TX + 2;
    constexpr int EH = TY + 2;
    constexpr int E = EW * EH;
    constexpr int NO = (TX * TY + NT - 1) / NT;
    constexpr int NTR = NTRC > 0 ? NTRC : 1;
    T* const s_u = lds.u;
    T* const s_v = lds.v;
    T* const s_eta = lds.Pot;
    T* const s_D = lds.D;
    T* const s_Pot = lds.Pot;
    T* const s_Ed = lds.Ed;
    T* const s_pll = lds.pll;
    T* const s_plt = lds.plt;
    auto& s_CH = lds.CH;
    auto& s_C = lds.C;
    uint8_t* const s_m = lds.m;

    CStatics<T>& a = *(CStatics<T>*)d.s;
    const int tid = threadIdx.x;
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

    constexpr int NE = (E + NT - 1) / NT;
    const auto& rm = R.m;
    const auto& ru = R.u;
    const auto& rv = R.v;
    const auto& re = R.e;
    const auto& rHo = R.Ho;
    const auto& rHl = R.Hl;
    const auto& rch = R.ch;
    const auto& rcb = R.cb;
    const T tz0 = R.tz0, tz1 = R.tz1, tcu = R.tcu, tcH = R.tcH, tfH = R.tfH, tcu0 = R.tcu0, tpc = R.tpc;
    const bool has_pc = a.mix && tid < EH * 24;
#pragma unroll
    for (int q = 0; q < NE; ++q) {
        const int e = tid + q * NT;
        if (e < E) {
            const int m = rm[q];
            const T eta = re[q];
            s_m[e] = (uint8_t)m;
            s_u[e] = ru[q];
            s_v[e] = rv[q];
            s_eta[e] = eta;
            if constexpr (VAR == VAR_MS) lds.et[e] = eta;
            const T He = OE(m) ? rHo[q] : rHl[q];
            T D;
            if (FN) D = OE(m) ? (eta + He) : T(1);   // computeD: D = eta + H on ocean eta, 1 elsewhere
            else D = He;                              // linear variants: D aliases H_eta
            s_D[e] = D;
#pragma unroll
            for (int n = 0; n < NTRC; ++n) {
                s_CH[n][e] = rch[n][q];
                // C = CH / h (TRC_tracer_advance), except the host-held C right after a hand-over
                s_C[n][e] = ((d.cbuf >> n) & 1) ? rcb[n][q] : dv(rch[n][q], D);
            }
        }
    }
    if (tid < EH) {
        lds.za[tid][0] = tz0;
        lds.za[tid][1] = tz1;
        lds.cu[tid + 1] = tcu;
        lds.cH[tid] = tcH;
        lds.rcH[tid] = T(1) / tcH;   // correctly rounded (dvr)
        lds.fH[tid] = tfH;
    }
    if (tid == NT - 1) lds.cu[0] = tcu0;
    if (has_pc) {
        const int r = tid / 24, c = tid % 24;
        if (c < 12) lds.pltc[r][c] = tpc;
        else lds.pllc[r][c - 12] = tpc;
    }
    if (EH * 24 > NT && a.mix)   // tiles taller than 8 rows: the remaining table rows
        for (int q = tid + NT; q < EH * 24; q += NT) {
            const int r = q / 24, c = q % 24;
            const int gj = wrap(j0 + r - 1, Ny);
            if (c < 12) lds.pltc[r][c] = gld(a.pltc, gj * 12 + c);
            else lds.pllc[r][c - 12] = gld(a.pllc, gj * 12 + c - 12);
        }
    using RTC = RowTabC<TX, TY, NTRC>;
    if constexpr (!RTC::UROW) {
#pragma unroll
        for (int q = 0; q < RTC::NQ; ++q) {
            const int x = tid + q * NT;
            if (x < RTC::R * RTC::K) lds.rt[x / RTC::K][x % RTC::K] = R.rt[q];
        }
    }
}

// Phases 2 and 3 of tile t, after its LDS stage and a barrier: derived fields,
// tendencies, AB integration and stores.  pg/pt: the outputs' AB history.
// OCEAN: every cell of the tile and of its halo is ocean at its u, v, eta and
// H points.  The masks are then the constant 15 and the mask arithmetic
// (no-land weights, land selects, coefficient rows) folds away at compile
// time; the expressions left are the general path's with every mask 1, so
// both give the same bits on such a tile.
template <typename T, int VAR, int NTRC, int TX, int TY, bool DIAG, bool DYN = true, bool OCEAN = false, typename L,
          typename G, typename GT, typename GE, typename LO>
__device__ __forceinline__ void tile_compute(const Dyn<T>& d, const int t, L& lds, const G& pg, const GT& pt,
                                             const GE& pe, const LO& plo) {
    constexpr int NT = 256;
    constexpr int EW = TX + 2;
    constexpr int EH = TY + 2;
    constexpr int E = EW * EH;
    constexpr int NO = (TX * TY + NT - 1) / NT;
    constexpr int NTR = NTRC > 0 ? NTRC : 1;
    T* const s_u = lds.u;
    T* const s_v = lds.v;
    T* const s_eta = lds.Pot;
    T* const s_D = lds.D;
    T* const s_Pot = lds.Pot;
    T* const s_Ed = lds.Ed;
    T* const s_pll = lds.pll;
    T* const s_plt = lds.plt;
    auto& s_CH = lds.CH;
    auto& s_C = lds.C;
    const uint8_t* const s_m = lds.m;
    // mask byte of LDS cell e
    auto M = [&](int e) -> uint8_t {
        if constexpr (OCEAN) return 15;
        else return s_m[e];
    };

    CStatics<T>& a = *(CStatics<T>*)d.s;
    const int tid = threadIdx.x;
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

    // ---- phase 2: derived fields on the extended tile
    constexpr bool SPLIT2 = !DIAG && !LdsDuDv<T, TX, TY>::SDU && !LdsDuDv<T, TX, TY>::SDV;
    if constexpr (SPLIT2) {
        // Two loops over the two regions the stencils need, instead of one loop
        // over the whole (TX+2)(TY+2) tile: EDens and P_lambda-lambda on
        // lx 0..TX, ly 0..TY, then (Pot overwrites eta in place) Dh, zeta, Pot
        // and P_lambda-theta on lx 1..TX+1, ly 1..TY+1.  Each has
        // (TX+1)(TY+1) cells; their second passes run on different waves
        // (tid ^ 128: wave 2 for the first loop, wave 0 for the second; wave 3
        // for the second measured 0.5 % slower).
        constexpr int RW = TX + 1, NR = (TX + 1) * (TY + 1);
        for (int r = tid ^ 128; r < NR; r += NT) {
            const int lx = r % RW, ly = r / RW;
            const int e = ly * EW + lx;
            const uint8_t m = M(e);
            const int ke = e + 1, kn = e + EW;
            T Ed = T(0);
            if (OE(m)) {
                if (FN) {
                    const T ue = s_u[ke], uc = s_u[e], vn = s_v[kn], vc = s_v[e];
                    Ed = fm(a.G, s_eta[e],
                            (fm(T(0.5), uc * uc, T(0.5) * (ue * ue)) + fm(T(0.5), vc * vc, T(0.5) * (vn * vn))) / T(2));
                } else if (VAR == VAR_MS) {   // g*eta + uU + vV
                    const int gj = wrap(j0 + ly - 1, Ny);
                    const int gi = wrap(i0 + lx - 1, Nx), gip = gi + 1 == Nx ? 0 : gi + 1;
                    const int gjp = gj + 1 == Ny ? 0 : gj + 1;
                    const T ube = fldg(a.ubs, (size_t)gj * P + gip, gj, OU(M(ke)) != 0);
                    const T ubc = fldg(a.ubs, (size_t)gj * P + gi, gj, OU(m) != 0);
                    const T vbn = fldg(a.vbs, (size_t)gjp * P + gi, gjp, OV(M(kn)) != 0);
                    const T vbc = fldg(a.vbs, (size_t)gj * P + gi, gj, OV(m) != 0);
                    const T ue = s_u[ke], uc = s_u[e], vn = s_v[kn], vc = s_v[e];
                    Ed = ((T(0.5) * (ue * ube) + T(0.5) * (uc * ubc)) + (T(0.5) * (vn * vbn) + T(0.5) * (vc * vbc))) +
                         a.G * s_eta[e];
                } else {
                    Ed = a.G * s_eta[e];
                }
            }
            s_Ed[e] = Ed;
            T pll = T(0);
            if (a.mix && OE(m)) {
                const T* c = &lds.pllc[ly][(OV(M(kn)) * 2 + OV(m)) * 3];
                pll = s_D[e] * fm(c[2], s_v[e], fm(c[1], s_v[kn], c[0] * (s_u[ke] - s_u[e])));
            }
            s_pll[e] = pll;
        }
        __syncthreads();   // every eta read before Pot takes its place
        for (int r = tid; r < NR; r += NT) {
            const int lx = r % RW + 1, ly = r / RW + 1;
            const int e = ly * EW + lx;
            const int gj = wrap(j0 + ly - 1, Ny);
            const uint8_t m = M(e);
            const int kw = e - 1, ks = e - EW, ksw = e - EW - 1;
            const int m1 = OE(m), m2 = OE(M(ksw)), m3 = OE(M(kw)), m4 = OE(M(ks));
            const T w = noland_w<T>(m1 + m2 + m3 + m4);
            const T Dh = fm(m4 ? w : T(0), s_D[ks],
                            fm(m3 ? w : T(0), s_D[kw], fm(m2 ? w : T(0), s_D[ksw], (m1 ? w : T(0)) * s_D[e])));
            T zeta = T(0);
            if (FN || VAR == VAR_MS) {
                T za_ = T(0), zb = T(0);
                if (OV(M(kw)) + OV(m) != 0)
                    za_ = lds.za[ly][OH(m) ? 0 : 1] * (s_v[e] - s_v[kw]);
                if (OU(m) + OU(M(ks)) != 0) {
                    const T bcH = OH(m) ? T(1) : a.lbc;
                    zb = dvr(bcH * fm(-lds.cu[ly], s_u[ks], lds.cu[ly + 1] * s_u[e]), a.AdT, a.rAdT);
                }
                zeta = za_ - dvr(zb, lds.cH[ly], lds.rcH[ly]);
            }
            const T pD = Dh > a.eps ? Dh : a.eps;
            T Pot;
            if constexpr (VAR == VAR_MS) {   // (f + zeta_bs) / pD
                const int gi = wrap(i0 + lx - 1, Nx);
                Pot = dv(lds.fH[ly] + fldg(a.zbs, (size_t)gj * P + gi, gj, OH(m)), pD);
                lds.zt[e] = zeta;
            } else {
                Pot = FN ? dv(lds.fH[ly] + zeta, pD) : dv(lds.fH[ly], pD);
            }
            s_Pot[e] = Pot;
            T plt = T(0);
            if (a.mix && OH(m)) {
                const T* c = &lds.pltc[ly][(OU(m) * 2 + OU(M(ks))) * 3];
                plt = Dh * fm(c[2], s_u[ks], fm(c[1], s_u[e], c[0] * (s_v[e] - s_v[kw])));
            }
            s_plt[e] = plt;
        }
    } else
    for (int e = tid; e < E; e += NT) {
        const int lx = e % EW, ly = e / EW;
        const int gj = wrap(j0 + ly - 1, Ny);
        const uint8_t m = M(e);
        const bool interior = DIAG && lx >= 1 && lx <= TX && ly >= 1 && ly <= TY &&
                              (i0 + lx - 1) < Nx && (j0 + ly - 1) < Ny;
        const size_t doff = interior ? (size_t)(j0 + ly - 1) * P + (i0 + lx - 1) : 0;
        const T eta_e = s_eta[e];   // before this cell's Pot takes its place
        if (lx >= 1 && ly >= 1) {
            const int kw = e - 1, ks = e - EW, ksw = e - EW - 1;
            // Dh = interp4(D, eta2H_noland): points (i,j), (i-1,j-1), (i-1,j), (i,j-1)
            const int m1 = OE(m), m2 = OE(M(ksw)), m3 = OE(M(kw)), m4 = OE(M(ks));
            const T w = noland_w<T>(m1 + m2 + m3 + m4);
            const T Dh = fm(m4 ? w : T(0), s_D[ks],
                            fm(m3 ? w : T(0), s_D[kw], fm(m2 ? w : T(0), s_D[ksw], (m1 ? w : T(0)) * s_D[e])));
            // computeRelVort -> vorticityFromVelocities2D (src/calc_lib.f90:663-678,
            // pder_zonal3D 892-926, pder_meridional3D 965-998)
            T zeta = T(0);
            if (FN || VAR == VAR_MS) {
                const int gjm = gj == 0 ? Ny - 1 : gj - 1;
                T za_ = T(0), zb = T(0);
                if (OV(M(kw)) + OV(m) != 0)
                    za_ = lds.za[ly][OH(m) ? 0 : 1] * (s_v[e] - s_v[kw]);
                if (OU(m) + OU(M(ks)) != 0) {
                    const T bcH = OH(m) ? T(1) : a.lbc;
                    zb = dvr(bcH * fm(-lds.cu[ly], s_u[ks], lds.cu[ly + 1] * s_u[e]), a.AdT, a.rAdT);
                }
                zeta = za_ - dvr(zb, lds.cH[ly], lds.rcH[ly]);
            }
            // computePotVort (src/swm_timestep_module.f90:230-257)
            const T pD = Dh > a.eps ? Dh : a.eps;
            T Pot;
            if constexpr (VAR == VAR_MS) {   // (f + zeta_bs) / pD
                const int gi = wrap(i0 + lx - 1, Nx);
                Pot = dv(lds.fH[ly] + fldg(a.zbs, (size_t)gj * P + gi, gj, OH(m)), pD);
                lds.zt[e] = zeta;
            } else {
                Pot = FN ? dv(lds.fH[ly] + zeta, pD) : dv(lds.fH[ly], pD);
            }
            s_Pot[e] = Pot;
            // SWM_LateralMixing_compute_p, P_lambda-theta on ocean H points
            T plt = T(0);
            if (a.mix && OH(m)) {
                const T* c = &lds.pltc[ly][(OU(m) * 2 + OU(M(ks))) * 3];
                plt = Dh * fm(c[2], s_u[ks], fm(c[1], s_u[e], c[0] * (s_v[e] - s_v[kw])));
            }
            s_plt[e] = plt;
            if (interior) { gst(a.diag[1], doff, Dh); gst(a.diag[4], doff, zeta); gst(a.diag[5], doff, Pot); }
        }
        // computeD: Du, Dv kept for phase 3 (Du from column 1, Dv from row 1 on)
        if constexpr (LdsDuDv<T, TX, TY>::SDU)
            if (lx >= 1) lds.Du[e] = interp_noland2(s_D[e], s_D[e - 1], OE(m), OE(M(e - 1)));
        if constexpr (LdsDuDv<T, TX, TY>::SDV)
            if (ly >= 1) lds.Dv[e] = interp_noland2(s_D[e], s_D[e - EW], OE(m), OE(M(e - EW)));
        if (interior) {   // computeMassFluxes: MU = Du * u on ocean u, MV = Dv * v on ocean v
            const T Du = interp_noland2(s_D[e], s_D[e - 1], OE(m), OE(M(e - 1)));
            const T Dv = interp_noland2(s_D[e], s_D[e - EW], OE(m), OE(M(e - EW)));
            gst(a.diag[2], doff, Du);
            gst(a.diag[7], doff, OU(m) ? Du * s_u[e] : T(0));
            gst(a.diag[3], doff, Dv);
            gst(a.diag[8], doff, OV(m) ? Dv * s_v[e] : T(0));
        }
        if (lx <= TX && ly <= TY) {
            const int ke = e + 1, kn = e + EW;
            // computeEdens (src/swm_timestep_module.f90:259-310); land keeps 0
            T Ed = T(0);
            if (OE(m)) {
                if (FN) {
                    const T ue = s_u[ke], uc = s_u[e], vn = s_v[kn], vc = s_v[e];
                    Ed = fm(a.G, eta_e,
                            (fm(T(0.5), uc * uc, T(0.5) * (ue * ue)) + fm(T(0.5), vc * vc, T(0.5) * (vn * vn))) / T(2));
                } else if (VAR == VAR_MS) {   // g*eta + uU + vV
                    const int gi = wrap(i0 + lx - 1, Nx), gip = gi + 1 == Nx ? 0 : gi + 1;
                    const int gjp = gj + 1 == Ny ? 0 : gj + 1;
                    const T ube = fldg(a.ubs, (size_t)gj * P + gip, gj, OU(M(ke)) != 0);
                    const T ubc = fldg(a.ubs, (size_t)gj * P + gi, gj, OU(m) != 0);
                    const T vbn = fldg(a.vbs, (size_t)gjp * P + gi, gjp, OV(M(kn)) != 0);
                    const T vbc = fldg(a.vbs, (size_t)gj * P + gi, gj, OV(m) != 0);
                    const T ue = s_u[ke], uc = s_u[e], vn = s_v[kn], vc = s_v[e];
                    Ed = ((T(0.5) * (ue * ube) + T(0.5) * (uc * ubc)) + (T(0.5) * (vn * vbn) + T(0.5) * (vc * vbc))) +
                         a.G * eta_e;
                } else {
                    Ed = a.G * eta_e;
                }
            }
            s_Ed[e] = Ed;
            // P_lambda-lambda on ocean eta points
            T pll = T(0);
            if (a.mix && OE(m)) {
                const T* c = &lds.pllc[ly][(OV(M(kn)) * 2 + OV(m)) * 3];
                pll = s_D[e] * fm(c[2], s_v[e], fm(c[1], s_v[kn], c[0] * (s_u[ke] - s_u[e])));
            }
            s_pll[e] = pll;
            if (interior) { gst(a.diag[6], doff, Ed); gst(a.diag[0], doff, s_D[e]); }
        }
    }
    __syncthreads();

    // ---- phase 3: tendencies, AB integration, clamp, tracers
    // With 64-wide tiles every wavefront owns one output row: the row index is
    // uniform and the per-row tables are read with scalar loads.
    constexpr bool UROW = (TX == 64 && NO == 1);
#pragma unroll
    for (int q = 0; q < NO; ++q) {
        const int o = tid + q * NT;
        const int lx = o % TX + 1, ly = o / TX + 1;
        const int i = i0 + lx - 1;
        const int j = UROW ? __builtin_amdgcn_readfirstlane(j0 + ly - 1) : j0 + ly - 1;
        if (o >= TX * TY || i >= Nx || j >= Ny) continue;
        const size_t off = (size_t)j * P + i;
        const int k = ly * EW + lx;
        const int ke = k + 1, kw = k - 1, kn = k + EW, ks = k - EW, knw = kn - 1, kse = ks + 1;
        const int jp = j + 1 == Ny ? 0 : j + 1;
        static_assert(RowTabC<TX, TY, NTRC>::UROW == UROW, "row-table layout");
        auto rv = [&](int e, int dr) -> T {   // rowtab entry e of row j (dr = 0) or jp (dr = 1)
            if constexpr (UROW) return ldc(a.rowtab, (size_t)(dr ? jp : j) * RT_N + e);
            else return lds.rt[ly - 1 + dr][e];
        };
        // rowtab entry (c ? e0 : e1) of row j: one LDS read at the selected
        // index, or both scalar loads and a select
        auto rvs = [&](bool c, int e0, int e1) -> T {
            if constexpr (UROW) {
                const T v0 = rv(e0, 0), v1 = rv(e1, 0);
                return c ? v0 : v1;
            } else {
                return lds.rt[ly - 1][c ? e0 : e1];
            }
        };
        // static field at (i, j); e = its rowtab entry, -1 for fields without one
        auto fl = [&](auto& f, int e, size_t at, bool ocean) -> T {
            if (f.full) return gld(f.full, at);
            if (e < 0) return gld(f.rows, 2 * (size_t)j + (ocean ? 0 : 1));
            return rvs(ocean, e, e + 1);
        };
        // x / (static field at (i, j)); e, er: its rowtab entry and that of its reciprocal
        auto flq = [&](T x, auto& f, int e, int er, size_t at, bool ocean) -> T {
            if (f.full) return dv(x, gld(f.full, at));
            return dvr(x, rvs(ocean, e, e + 1), rvs(ocean, er, er + 1));
        };
        // TRC_coeff magnitude (trc_adv) from the rowtab pair at e
        auto ta = [&](bool neg, int nummask, int den, int e) -> T {
            const T v = rvs(den == 1, e, e + 1);
            if (den == 0) return __builtin_nan("");
            if (!nummask) return neg ? T(-0.0) : T(0.0);
            return neg ? -v : v;
        };
        const uint8_t m = M(k);
        const int oe = OE(m), ou = OU(m), ov = OV(m), oh = OH(m);
        constexpr bool SDU = LdsDuDv<T, TX, TY>::SDU, SDV = LdsDuDv<T, TX, TY>::SDV;
        // Du, Dv at a halo cell: from LDS, or interp2(D, eta2u/eta2v_noland) again
        auto DUat = [&](int kk) -> T {
            if constexpr (SDU) return lds.Du[kk];
            else return interp_noland2(s_D[kk], s_D[kk - 1], OE(M(kk)), OE(M(kk - 1)));
        };
        auto DVat = [&](int kk) -> T {
            if constexpr (SDV) return lds.Dv[kk];
            else return interp_noland2(s_D[kk], s_D[kk - EW], OE(M(kk)), OE(M(kk - EW)));
        };
        const T Du = DUat(k);
        const T Dv = DVat(k);
        // computeMassFluxes at the points the stencils need
        auto MUat = [&](int kk) -> T { return OU(M(kk)) ? DUat(kk) * s_u[kk] : T(0); };
        auto MVat = [&](int kk) -> T { return OV(M(kk)) ? DVat(kk) * s_v[kk] : T(0); };
        const T MUk = ou ? Du * s_u[k] : T(0), MVk = ov ? Dv * s_v[k] : T(0);
        const T MUe = MUat(ke), MUs = MUat(ks), MUse = MUat(kse);
        const T MVn = MVat(kn), MVw = MVat(kw), MVnw = MVat(knw);

        // SWM_forcing_update (src/swm_forcing_module.f90:135-170)
        T Fe = fl(a.Fec, RT_FEC, off, oe);
        if (a.has_oe) Fe = Fe + (T(2) * (fl(a.oRe, -1, off, oe) * d.oc[2] - fl(a.oIe, -1, off, oe) * d.os[2])) * T(oe);
        T Fx = fl(a.Fxc, RT_FXC, off, ou);
        if (a.has_taux && ou) {
            const T tx = fl(a.taux, RT_TAUX, off, true);
            Fx = Fx + dv(a.has_tau_scale ? a.tau_scale * tx : tx, a.RHO0 * Du);
        }
        if (a.has_ox) Fx = Fx + (T(2) * (fl(a.oRx, -1, off, ou) * d.oc[0] - fl(a.oIx, -1, off, ou) * d.os[0])) * T(ou);
        T Fy = fl(a.Fyc, RT_FYC, off, ov);
        if (a.has_tauy && ov) {
            const T ty = fl(a.tauy, RT_TAUY, off, true);
            Fy = Fy + dv(a.has_tau_scale ? a.tau_scale * ty : ty, a.RHO0 * Dv);
        }
        if (a.has_oy) Fy = Fy + (T(2) * (fl(a.oRy, -1, off, ov) * d.oc[1] - fl(a.oIy, -1, off, ov) * d.os[1])) * T(ov);
        if (a.nstream) apply_streams<T>(a, d, off, ou, ov, oe, Du, Dv, Fx, Fy, Fe);

        // SWM_timestep_nonlinear (src/swm_timestep_module.f90:349-443)
        T Ge = T(0), Gu = T(0), Gv = T(0), Lu = T(0), Lv = T(0);
        // neighbours' offsets for the basic-state terms (LINEARISED_MEAN_STATE)
        const int ipx = i + 1 == Nx ? 0 : i + 1, imx = i == 0 ? Nx - 1 : i - 1, jmx = j == 0 ? Ny - 1 : j - 1;
        if (oe) {
            const T t1 = dvr(-(MUe - MUk), a.dL, a.rdL);
            const T t2 = dvr(fm(-MVk, rv(RT_CV, 0), MVn * rv(RT_CV, 1)), a.dT, a.rdT);
            T tt = t1 - t2;
            if constexpr (VAR == VAR_MS) {
                // src/swm_timestep_module.f90:359-365.  The first interpolation is
                // taken at i index ip1(j) (column j+1 of row j), as the reference
                // indexes it; rows j < Nx only (create rejects Ny > Nx).
                const int c = j + 1 == Nx ? 0 : j + 1, cm = c == 0 ? Nx - 1 : c - 1;
                const size_t rc = (size_t)j * P;
                const T E1 = interp_noland2(gld(se, rc + c), gld(se, rc + cm), OE(gld(a.mask, rc + c)),
                                            OE(gld(a.mask, rc + cm)));
                const T E0 = interp_noland2(lds.et[k], lds.et[kw], oe, OE(M(kw)));
                const T ube = fldg(a.ubs, (size_t)j * P + ipx, j, OU(M(ke)) != 0);
                const T ubk = fldg(a.ubs, off, j, ou != 0);
                tt = tt - dvr(E1 * ube - E0 * ubk, a.dL, a.rdL);
                const T E2 = interp_noland2(lds.et[kn], lds.et[k], OE(M(kn)), oe);
                const T E3 = interp_noland2(lds.et[k], lds.et[ks], oe, OE(M(ks)));
                const T vbn = fldg(a.vbs, (size_t)jp * P + i, jp, OV(M(kn)) != 0);
                const T vbk = fldg(a.vbs, off, j, ov != 0);
                tt = tt - dvr((rv(RT_CV, 1) * E2) * vbn - (rv(RT_CV, 0) * E3) * vbk, a.dT, a.rdT);
            }
            Ge = dvr(tt, rv(RT_ACE, 0), rv(RT_RACE, 0)) + Fe;
        }
        if (ou) {
            const T P2 = fm(T(0.5), s_Pot[k], T(0.5) * s_Pot[kn]);
            const T M4 = fm(T(0.25), MVk, fm(T(0.25), MVnw, fm(T(0.25), MVw, T(0.25) * MVn)));
            T g = P2 * M4;
            if constexpr (VAR == VAR_MS) {   // + interp2(zeta, H2u) * interp4(v_bs, v2u)
                const T Z2 = T(0.5) * lds.zt[kn] + T(0.5) * lds.zt[k];
                const T vbn = fldg(a.vbs, (size_t)jp * P + i, jp, OV(M(kn)) != 0);
                const T vbw = fldg(a.vbs, (size_t)j * P + imx, j, OV(M(kw)) != 0);
                const T vbnw = fldg(a.vbs, (size_t)jp * P + imx, jp, OV(M(knw)) != 0);
                const T vbk = fldg(a.vbs, off, j, ov != 0);
                g = g + Z2 * (((T(0.25) * vbn + T(0.25) * vbw) + T(0.25) * vbnw) + T(0.25) * vbk);
            }
            g = g - dvr(s_Ed[k] - s_Ed[kw], rv(RT_EDU, 0), rv(RT_REDU, 0));
            if (a.quad) {
                const T vi = ((T(0.25) * s_v[kn] + T(0.25) * s_v[kw]) + T(0.25) * s_v[knw]) + T(0.25) * s_v[k];
                const T uc = s_u[k];
                T qf = fl(a.gsu, -1, off, true) * sqrt(uc * uc + (vi * vi)) * uc;
                if (a.barotropic) qf = dv(qf, fl(a.Hu, -1, off, true));
                g = g - qf;
            }
            if (a.mix) {
                const int ot = (oh + OH(M(kn))) < 1 ? 1 : (oh + OH(M(kn)));
                const T c0 = rvs(ot == 1, RT_LMU, RT_LMU + 3), c1 = rvs(ot == 1, RT_LMU + 1, RT_LMU + 4),
                        c2 = rvs(ot == 1, RT_LMU + 2, RT_LMU + 5);
                Lu = dv(fm(c2, s_plt[k], fm(c1, s_plt[kn], c0 * (s_pll[k] - s_pll[kw]))), Du);
                g = g + Lu;
            }
            Gu = g + Fx;
        }
        if (ov) {
            const T P2 = fm(T(0.5), s_Pot[k], T(0.5) * s_Pot[ke]);
            const T M4 = fm(T(0.25), MUse, fm(T(0.25), MUk, fm(T(0.25), MUs, T(0.25) * MUe)));
            T g = -(P2 * M4);
            if constexpr (VAR == VAR_MS) {   // - interp2(zeta, H2v) * interp4(u_bs, u2v)
                const T Z2 = T(0.5) * lds.zt[ke] + T(0.5) * lds.zt[k];
                const T ube = fldg(a.ubs, (size_t)j * P + ipx, j, OU(M(ke)) != 0);
                const T ubs_ = fldg(a.ubs, (size_t)jmx * P + i, jmx, OU(M(ks)) != 0);
                const T ubk = fldg(a.ubs, off, j, ou != 0);
                const T ubse = fldg(a.ubs, (size_t)jmx * P + ipx, jmx, OU(M(kse)) != 0);
                g = g - Z2 * (((T(0.25) * ube + T(0.25) * ubs_) + T(0.25) * ubk) + T(0.25) * ubse);
            }
            g = g - dvr(s_Ed[k] - s_Ed[ks], a.AdT, a.rAdT);
            if (a.quad) {
                const T ui = ((T(0.25) * s_u[ke] + T(0.25) * s_u[ks]) + T(0.25) * s_u[k]) + T(0.25) * s_u[kse];
                const T vc = s_v[k];
                T qf = fl(a.gsv, -1, off, true) * sqrt(vc * vc + (ui * ui)) * vc;
                if (a.barotropic) qf = dv(qf, fl(a.Hv, -1, off, true));
                g = g - qf;
            }
            if (a.mix) {
                const int ot = (oe + OE(M(ks))) < 1 ? 1 : (oe + OE(M(ks)));
                const T c0 = rvs(ot == 1, RT_LMV, RT_LMV + 3), c1 = rvs(ot == 1, RT_LMV + 1, RT_LMV + 4),
                        c2 = rvs(ot == 1, RT_LMV + 2, RT_LMV + 5);
                Lv = dv(fm(c2, s_pll[ks], fm(c1, s_pll[k], c0 * (s_plt[ke] - s_plt[k]))), Dv);
                g = g + Lv;
            }
            Gv = g + Fy;
        }

        if (DIAG) {
            gst(a.diag[9], off, Fx); gst(a.diag[10], off, Fy); gst(a.diag[11], off, Fe);
            gst(a.diag[12], off, Lu); gst(a.diag[13], off, Lv);
#pragma unroll
            for (int n = 0; n < NTRC; ++n) {
                const T C = s_C[n][k];
                gst(a.tdiag[n][0], off, C);
                gst(a.tdiag[n][1], off, MUk * T(0.5) * (C + s_C[n][kw]));
                gst(a.tdiag[n][2], off, MVk * T(0.5) * (C + s_C[n][ks]));
            }
            continue;
        }

        // integrate_AB (src/time_integration_module.f90:74-112); land keeps the
        // previous newest tendency, which the reference never overwrites there.
        const T Gu3 = ou ? Gu : pg[q][1];
        const T Gv3 = ov ? Gv : pg[q][3];
        const T Ge3 = oe ? Ge : pg[q][5];
        // the damping static at (i, j): FP32 contexts hold 1 - 1/impl there
        auto fdv = [&](auto& f, int e, bool ocean) -> T { return f.full ? gld(f.full, off) : rvs(ocean, e, e + 1); };
        T un, vn, en;
        unsigned short lu = 0, lv = 0, le = 0;
        if constexpr (sizeof(T) == 4) {   // increments added to the carried pair
            T iu = d.dtc1 * pg[q][0], iv = d.dtc1 * pg[q][2], ie = d.dtc1 * pg[q][4];
            if (d.tstep >= 2) {
                iu = madd(iu, d.dtc2, pg[q][1]);
                iv = madd(iv, d.dtc2, pg[q][3]);
                ie = madd(ie, d.dtc2, pg[q][5]);
            }
            if (d.tstep >= 3) {
                iu = madd(iu, d.dtc3, Gu3);
                iv = madd(iv, d.dtc3, Gv3);
                ie = madd(ie, d.dtc3, Ge3);
            }
            un = carry_step(s_u[k], plo[q][0], iu, fdv(a.imu, RT_IMU, ou), lu);
            vn = carry_step(s_v[k], plo[q][1], iv, fdv(a.imv, RT_IMV, ov), lv);
            en = carry_step(pe[q], plo[q][2], ie, fdv(a.ime, RT_IME, oe), le);
        } else {
            un = s_u[k], vn = s_v[k], en = pe[q];
            un = madd(un, d.dtc1, pg[q][0]);
            vn = madd(vn, d.dtc1, pg[q][2]);
            en = madd(en, d.dtc1, pg[q][4]);
            if (d.tstep >= 2) {
                un = madd(un, d.dtc2, pg[q][1]);
                vn = madd(vn, d.dtc2, pg[q][3]);
                en = madd(en, d.dtc2, pg[q][5]);
            }
            if (d.tstep >= 3) {
                un = madd(un, d.dtc3, Gu3);
                vn = madd(vn, d.dtc3, Gv3);
                en = madd(en, d.dtc3, Ge3);
            }
            un = flq(un, a.imu, RT_IMU, RT_RIMU, off, ou);
            vn = flq(vn, a.imv, RT_IMV, RT_RIMV, off, ov);
            en = flq(en, a.ime, RT_IME, RT_RIME, off, oe);
        }
        if (FN) {   // SWM_timestep_step: where (eta < minD - H) eta = minD - H
            const T lim = a.minD - fl(a.He, RT_HE, off, oe);
            if (en < lim) { en = lim; le = 0; }
        }
        gst(st, (PL_U + 1 - d.cur) * pl + off, un);
        gst(st, (PL_V + 1 - d.cur) * pl + off, vn);
        gst(st, (PL_E + 1 - d.cur) * pl + off, en);
        if constexpr (sizeof(T) == 4) {
            if (a.carry) {
                gst((unsigned short*)lo_plane(st, pl, PL_LO), off, lu);
                gst((unsigned short*)lo_plane(st, pl, PL_LO + 1), off, lv);
                gst((unsigned short*)lo_plane(st, pl, PL_LO + 2), off, le);
            }
        }
        gst(st, (PL_GU + r2) * pl + off, Gu3);
        gst(st, (PL_GV + r2) * pl + off, Gv3);
        gst(st, (PL_GE + r2) * pl + off, Ge3);

        // TRC_tracer_incr + TRC_tracer_integrate (src/tracer_module.f90:154-232)
#pragma unroll
        for (int n = 0; n < NTRC; ++n) {
            T* const tb = st + (PL_TRC + PL_PER_TRC * n) * pl;
            const T CH = s_CH[n][k], CHe = s_CH[n][ke], CHw = s_CH[n][kw], CHn = s_CH[n][kn], CHs = s_CH[n][ks];
            const T C = s_C[n][k], Ce = s_C[n][ke], Cw = s_C[n][kw], Cn = s_C[n][kn], Cs = s_C[n][ks];
            // tracer velocities are the master fields (model_advance,
            // src/model.f90:238-259): u = 0, += ocean_u * DYNFROMFILE data
            // (DFF_advance, src/dynFromFile_module.f90:137-160), += SWM_u
            T um = T(0) + s_u[k], ume = T(0) + s_u[ke], vm = T(0) + s_v[k], vmn = T(0) + s_v[kn];
            const uint8_t me = M(ke), mw = M(kw), mn = M(kn), ms = M(ks);
            if (DYN && a.dyn_u) {
                const size_t offe = (size_t)j * P + (i + 1 == Nx ? 0 : i + 1);
                um = (T(0) + T(ou) * dyn_val(a.dyn_u, d.ds[0], off)) + s_u[k];
                ume = (T(0) + T(OU(me)) * dyn_val(a.dyn_u, d.ds[0], offe)) + s_u[ke];
            }
            if (DYN && a.dyn_v) {
                const size_t offn = (size_t)jp * P + i;
                vm = (T(0) + T(ov) * dyn_val(a.dyn_v, d.ds[1], off)) + s_v[k];
                vmn = (T(0) + T(OV(mn)) * dyn_val(a.dyn_v, d.ds[1], offn)) + s_v[kn];
            }
            const int oee = OE(me), oew = OE(mw), oen = OE(mn), oes = OE(ms);
            const int oue = OU(me), ovn = OV(mn);
            T diff = T(0);
            if (oe) {
                const T kap = a.kappa[n];
                const T hue = DUat(ke);   // interp2 of D(i+1,j), D(i,j)
                const T hvn = DVat(kn);   // interp2 of D(i,j+1), D(i,j)
                const T c9 = oue ? rv(RT_T9, 0) : T(0.0);
                const T c10 = ou ? -rv(RT_T9, 0) : T(-0.0);
                const T c11 = ovn ? rv(RT_T11, 0) : T(0.0);
                const T c12 = ov ? -rv(RT_T12, 0) : T(-0.0);
                diff = fm((kap * Dv) * (C - Cs), c12,
                          fm((kap * hvn) * (Cn - C), c11, fm((kap * Du) * (C - Cw), c10, ((kap * hue) * (Ce - C)) * c9)));
            }
            const T c1 = ta(true, oue & oee, oee + oe, RT_TA);
            const T c2 = ta(false, ou & oew, oew + oe, RT_TA);
            const T c3 = ta(true, ovn & oen, oen + oe, RT_TB);
            const T c4 = ta(false, ov & oes, oes + oe, RT_TC);
            const T c5 = ta(true, oue & oe, oee + oe, RT_TA);
            const T c6 = ta(false, ou & oe, oew + oe, RT_TA);
            const T c7 = ta(true, ovn & oe, oen + oe, RT_TB);
            const T c8 = ta(false, ov & oe, oes + oe, RT_TC);
            T g = fm(CHw * um, c2, (CHe * ume) * c1);
            g = fm(CHn * vmn, c3, g);
            g = fm(CHs * vm, c4, g);
            g = fm(CH * ume, c5, g);
            g = fm(CH * um, c6, g);
            g = fm(CH * vmn, c7, g);
            g = fm(CH * vm, c8, g);
            g = g + diff;
            g = fm(-(fl(a.gC[n], RT_TRC + RT_PER_TRC * n + 2, off, oe) * s_D[k]),
                   C - fl(a.C0[n], RT_TRC + RT_PER_TRC * n + 4, off, oe), g);
            g = fm(C, Fe, g);   // + C * Fe
            T chn;
            if constexpr (sizeof(T) == 4) {
                T inc = d.dtc1 * pt[q][n][0];
                if (d.tstep >= 2) inc = madd(inc, d.dtc2, pt[q][n][1]);
                if (d.tstep >= 3) inc = madd(inc, d.dtc3, g);
                unsigned short lc;
                chn = carry_step(CH, plo[q][3 + n], inc, fdv(a.timpl[n], RT_TRC + RT_PER_TRC * n, oe), lc);
                if (a.carry) gst((unsigned short*)lo_plane(tb, pl, PL_TLO), off, lc);
            } else {
                chn = CH + d.dtc1 * pt[q][n][0];
                if (d.tstep >= 2) chn = madd(chn, d.dtc2, pt[q][n][1]);
                if (d.tstep >= 3) chn = madd(chn, d.dtc3, g);
                chn = flq(chn, a.timpl[n], RT_TRC + RT_PER_TRC * n, RT_TRC + RT_PER_TRC * n + 6, off, oe);
            }
            gst(tb, (1 - d.cur) * pl + off, chn);
            gst(tb, (2 + r2) * pl + off, g);
        }
    }
}

// Whether the step and batched kernels split off all-ocean tiles (OSPLIT of
// tile_body): FP32 only.  On the 1024^2 basin with 5 % islands 70 % of the
// 32x8 tiles are all ocean, and the FP32 step is 6 % faster (31.7 -> 29.8 us).
// FP64 is 3.7 % slower with it: the two instantiations in one kernel raise its
// VGPRs from 93 to 107 (four waves per SIMD instead of five), which costs more
// than the mask arithmetic saves (7 % with every tile on the OCEAN path).
template <typename T>
inline constexpr bool kOceanSplit = sizeof(T) == 4;

// One 256-thread block advances tile t (TX x TY cells) of the domain in d by
// one step.  L is a TileLds with room for at least NTRC tracers.
// LATE: issue the outputs' AB-history loads after the halo is in LDS instead
// of with the halo loads.  When every block of the grid starts at once (a few
// rounds of resident blocks, as at 1024^2) all blocks load together and then
// compute together; moving a third of the loads behind the halo stage spreads
// that burst over the barrier and phase 2 (-6 % at 1024^2).  With many rounds
// the blocks drift apart and the early prefetch hides more latency (+3 % at
// 2048^2 for the late order).
// DYN: the tracer may read DYNFROMFILE datasets (contexts that step alone;
// the batched kernel is built without that branch).
// OSPLIT: tiles whose cells and halo are all ocean take tile_compute's
// OCEAN instantiation.  Each wavefront votes over the masks of its halo cells
// (already in registers) and writes the vote to LDS before the barrier that
// follows the stage; no further barrier is needed.  The step and batched
// kernels set it for FP32 (kOceanSplit).
template <typename T, int VAR, int NTRC, int TX, int TY, bool DIAG, bool LATE = false, bool DYN = true,
          bool OSPLIT = false, typename L>
__device__ __forceinline__ void tile_body(const Dyn<T>& d, const int t, L& lds) {
    using S = TileShapeC<TX, TY>;
    TileRegs<T, S::NE, S::NO, (NTRC > 0 ? NTRC : 1), RowTabC<TX, TY, NTRC>::NQ> R;
    tile_load<T, VAR, NTRC, TX, TY, DIAG, !LATE>(d, t, R);
    tile_stage<T, VAR, NTRC, TX, TY, DIAG>(d, t, R, lds);
    if (LATE) tile_load_hist<T, VAR, NTRC, TX, TY, DIAG>(d, t, R);
    if constexpr (OSPLIT && !DIAG) {
        static_assert(S::NT == 256, "four wavefronts vote");
        const int w = __all(R.ocean);
        if ((threadIdx.x & 63) == 0) lds.ocean[threadIdx.x >> 6] = w;
        __syncthreads();
        if (lds.ocean[0] & lds.ocean[1] & lds.ocean[2] & lds.ocean[3])
            tile_compute<T, VAR, NTRC, TX, TY, DIAG, DYN, true>(d, t, lds, R.g, R.gt, R.eo, R.lo);
        else
            tile_compute<T, VAR, NTRC, TX, TY, DIAG, DYN, false>(d, t, lds, R.g, R.gt, R.eo, R.lo);
    } else {
        __syncthreads();
        tile_compute<T, VAR, NTRC, TX, TY, DIAG, DYN>(d, t, lds, R.g, R.gt, R.eo, R.lo);
    }
}

template <typename T, int VAR, int NTRC, int TX, int TY, bool DIAG, bool LATE = false>
__global__ void __launch_bounds__(256)
step_kernel(const Dyn<T> d) {
    using DD = LdsDuDv<T, TX, TY>;
    using RTC = RowTabC<TX, TY, NTRC>;
    __shared__ TileLds<T, (NTRC > 0 ? NTRC : 1), (TX + 2) * (TY + 2), TY + 2,
                       (VAR == VAR_MS ? (TX + 2) * (TY + 2) : 1), DD::DUE, DD::DVE, RTC::LR, RTC::LK> lds;
    tile_body<T, VAR, NTRC, TX, TY, DIAG, LATE, true, kOceanSplit<T>>(d, xcd_tile(blockIdx.x, d.nblk), lds);
}

// Persistent tile kernel: block b advances the tiles [b*tpb, (b+1)*tpb) in
// order (neighbours along x, so halo columns are shared in L2) and issues the
// global loads of tile k+1 before it computes tile k, so the memory latency of
// one tile overlaps the arithmetic of the previous one.
template <typename T, int VAR, int NTRC, int TX, int TY>
__global__ void __launch_bounds__(256)
step_kernel_p(const Dyn<T> d) {
    using S = TileShapeC<TX, TY>;
    constexpr int NTR = NTRC > 0 ? NTRC : 1;
    using DD = LdsDuDv<T, TX, TY>;
    using RTC = RowTabC<TX, TY, NTRC>;
    __shared__ TileLds<T, NTR, S::E, S::EH, 1, DD::DUE, DD::DVE, RTC::LR, RTC::LK> lds;
    const int b = xcd_tile(blockIdx.x, gridDim.x);
    const int t0 = b * d.tpb;
    const int t1 = t0 + d.tpb < d.nblk ? t0 + d.tpb : d.nblk;
    if (t0 >= t1) return;   // whole block
    TileRegs<T, S::NE, S::NO, NTR, RTC::NQ> R;
    tile_load<T, VAR, NTRC, TX, TY, false>(d, t0, R);
    for (int t = t0; t < t1; ++t) {
        if (t > t0) __syncthreads();   // phase 3 of the previous tile is done with LDS
        tile_stage<T, VAR, NTRC, TX, TY, false>(d, t, R, lds);
        T g[S::NO][6], gt[S::NO][NTR][2], eo[S::NO];
        unsigned short lo[S::NO][3 + NTR];
#pragma unroll
        for (int q = 0; q < S::NO; ++q) {
            eo[q] = R.eo[q];
#pragma unroll
            for (int k = 0; k < 3 + NTR; ++k) lo[q][k] = R.lo[q][k];
#pragma unroll
            for (int k = 0; k < 6; ++k) g[q][k] = R.g[q][k];
#pragma unroll
            for (int n = 0; n < NTR; ++n) { gt[q][n][0] = R.gt[q][n][0]; gt[q][n][1] = R.gt[q][n][1]; }
        }
        __syncthreads();
        if (t + 1 < t1) tile_load<T, VAR, NTRC, TX, TY, false>(d, t + 1, R);
        tile_compute<T, VAR, NTRC, TX, TY, false>(d, t, lds, g, gt, eo, lo);
    }
}

// ---------------------------------------------------------------------------
// Batched step: many independent domains of uneven size advance by one step in
// one launch.  Work item w is tile (w - wbase) of member owner[w]; the
// per-member step state (level, AB ring slot, itt, steps left) lives in
// device memory and the last block of each step moves it on, so a captured
// graph of the two launches replays any number of steps.
// ---------------------------------------------------------------------------
template <typename T>
struct Member {
    const Statics<T>* s;
    T* state;
    long long plane;
    int ntx;          // tile columns
    int ntile;        // tiles of this member
    int wbase;        // first work item of this member
    int ntr;          // tracers
    T dtc[3][3];      // dt*ab_coeff(ti, tstep) for tstep = 1..3
};

struct MemberStep {
    long long itt;        // steps done
    long long remaining;  // steps still to do
    int cur, ring, cbuf, pad;
};

// Work item w of a batched step: tile (w - wbase) of member owner[w].
template <typename T, int VAR, int NTRC_MAX, int TX, int TY, typename L>
__device__ __forceinline__ void batch_item(const Member<T>* __restrict__ mem, const MemberStep* __restrict__ stp,
                                           const int* __restrict__ owner, const int w, L& lds) {
    const int lo = owner[w];   // member whose work range holds w
    const Member<T>& m = mem[lo];
    const MemberStep ms = stp[lo];
    if (ms.remaining <= 0) return;   // whole block: the member is done
    Dyn<T> d{};
    d.s = m.s;
    d.state = m.state;
    d.plane = m.plane;
    d.ntx = m.ntx;
    d.nblk = m.ntile;
    d.ty = TY;
    d.cur = ms.cur;
    d.ring = ms.ring;
    d.cbuf = ms.cbuf;
    const long long itt = ms.itt + 1;
    const int tstep = itt < 3 ? (int)itt : 3;
    d.tstep = tstep;
    d.dtc1 = m.dtc[tstep - 1][0];
    d.dtc2 = m.dtc[tstep - 1][1];
    d.dtc3 = m.dtc[tstep - 1][2];
    const int t = w - m.wbase;
    if constexpr (NTRC_MAX == 0) {
        tile_body<T, VAR, 0, TX, TY, false, false, false, kOceanSplit<T>>(d, t, lds);
    } else if constexpr (NTRC_MAX == 1) {
        if (m.ntr == 0) tile_body<T, VAR, 0, TX, TY, false, false, false, kOceanSplit<T>>(d, t, lds);
        else tile_body<T, VAR, 1, TX, TY, false, false, false, kOceanSplit<T>>(d, t, lds);
    } else {
        if (m.ntr == 0) tile_body<T, VAR, 0, TX, TY, false, false, false, kOceanSplit<T>>(d, t, lds);
        else if (m.ntr == 1) tile_body<T, VAR, 1, TX, TY, false, false, false, kOceanSplit<T>>(d, t, lds);
        else tile_body<T, VAR, 2, TX, TY, false, false, false, kOceanSplit<T>>(d, t, lds);
    }
}

// model_advance of every member that stepped: N0 <- N0p1, G ring moves on.
__device__ __forceinline__ void advance_members(MemberStep* stp, int nmem) {
    for (int k = threadIdx.x; k < nmem; k += 256) {
        MemberStep s = stp[k];
        if (s.remaining <= 0) continue;
        s.itt += 1;
        s.remaining -= 1;
        s.cur = 1 - s.cur;
        s.ring = s.ring == 2 ? 0 : s.ring + 1;
        s.cbuf = 0;
        stp[k] = s;
    }
}

// The block that finishes last (device-scope counter ctr[kBatchQueues], fenced)
// advances every member and clears the step's counters, so one launch is a
// whole batched step: every other block has read its member's step state
// before it counted itself done.  Used by the persistent queue kernel, whose
// grid is the resident blocks; with one block per tile (tens of thousands of
// blocks) the atomics on one counter serialise and cost more than the
// separate batch_advance launch.
constexpr int kBatchQueues = 8;
__device__ __forceinline__ void batch_finish(MemberStep* stp, int nmem, int* ctr) {
    __shared__ int s_last;
    __syncthreads();   // the block's work is done
    if (threadIdx.x == 0) {
        __threadfence();
        s_last = atomicAdd(&ctr[kBatchQueues], 1) == (int)gridDim.x - 1;
    }
    __syncthreads();
    if (!s_last) return;
    __threadfence();
    advance_members(stp, nmem);
    if (threadIdx.x <= kBatchQueues) ctr[threadIdx.x] = 0;   // tile queues and this counter
}

// One block per work item; blocks of one XCD take a contiguous run of items
// (neighbouring tiles of a member share halo lines in that XCD's L2).
template <typename T, int VAR, int NTRC_MAX, int TX, int TY>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5)))
batch_kernel(const Member<T>* __restrict__ mem, const MemberStep* __restrict__ stp, const int* __restrict__ owner,
             int nwork) {
    using DD = LdsDuDv<T, TX, TY>;
    using RTC = RowTabC<TX, TY, NTRC_MAX>;
    __shared__ TileLds<T, (NTRC_MAX > 0 ? NTRC_MAX : 1), (TX + 2) * (TY + 2), TY + 2, 1, DD::DUE, DD::DVE, RTC::LR,
                       RTC::LK> lds;
    batch_item<T, VAR, NTRC_MAX, TX, TY>(mem, stp, owner, xcd_tile(blockIdx.x, nwork), lds);
}

// model_advance after a batch_kernel step (one thread per member)
__global__ void __launch_bounds__(256) batch_advance(MemberStep* stp, int nmem) {
    const int b0 = blockIdx.x * 256;   // this block's members: [b0, b0 + 256)
    if (b0 < nmem) advance_members(stp + b0, nmem - b0 < 256 ? nmem - b0 : 256);
}

// Persistent tile queue: a grid of resident blocks drains the step's work
// items through eight counters, one per XCD.  Queue x holds the contiguous run
// of items that xcd_tile gives XCD x; a block (dispatched to XCD blockIdx % 8)
// takes items from its own queue and, when that is empty, from the next ones
// in turn, so uneven members (tracers, tile counts) do not leave an XCD idle
// while another still has a tail.  A claim takes kQueueChunk items, and the
// block issues its next claim before it works through the current one, so the
// atomic's round trip overlaps the tiles.  Every block ends after finding all
// eight queues empty.  The last block to finish clears the counters for the
// next step (batch_finish), so a captured step graph holds no memset.
constexpr int kQueueChunk = 2;
template <typename T, int VAR, int NTRC_MAX, int TX, int TY>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4)))
batch_queue_kernel(const Member<T>* __restrict__ mem, MemberStep* __restrict__ stp,
                   const int* __restrict__ owner, int nwork, int nmem, int* __restrict__ queue) {
    using DD = LdsDuDv<T, TX, TY>;
    using RTC = RowTabC<TX, TY, NTRC_MAX>;
    __shared__ TileLds<T, (NTRC_MAX > 0 ? NTRC_MAX : 1), (TX + 2) * (TY + 2), TY + 2, 1, DD::DUE, DD::DVE, RTC::LR,
                       RTC::LK> lds;
    __shared__ int s_item;
    const int q0 = nwork >> 3, r0 = nwork & 7;
    int x = blockIdx.x & 7;
    int claim = 0;   // thread 0: the claim in flight on queue x
    if (threadIdx.x == 0) claim = atomicAdd(&queue[x], kQueueChunk);
    for (int empty = 0; empty < kBatchQueues;) {
        __syncthreads();   // the previous item is done with LDS and s_item
        if (threadIdx.x == 0) s_item = claim;
        __syncthreads();
        const int k = s_item;
        const int lo = x * q0 + (x < r0 ? x : r0);
        const int n = q0 + (x < r0 ? 1 : 0);
        if (k >= n) {   // uniform: this queue is drained; claim from the next one
            x = (x + 1) & 7;
            ++empty;
            if (threadIdx.x == 0 && empty < kBatchQueues) claim = atomicAdd(&queue[x], kQueueChunk);
            continue;
        }
        if (threadIdx.x == 0) claim = atomicAdd(&queue[x], kQueueChunk);   // next claim, in flight meanwhile
        const int kend = k + kQueueChunk < n ? k + kQueueChunk : n;
        for (int kk = k; kk < kend; ++kk) {
            if (kk > k) __syncthreads();   // the previous tile is done with LDS
            batch_item<T, VAR, NTRC_MAX, TX, TY>(mem, stp, owner, lo + kk, lds);
        }
    }
    batch_finish(stp, nmem, queue);
}


// ---------------------------------------------------------------------------
// Marching-wavefront step kernel.
//
// Each 64-lane wavefront owns a strip of 64 columns (62 outputs plus one halo
// lane on each side) and a segment of d.ty output rows.  It walks down its rows
// one per iteration, keeps the three rows j-1, j, j+1 of the state and of the
// derived fields in registers, takes horizontal neighbours with full-wavefront
// DPP shifts, and loads the next row and the AB history one iteration ahead of
// use.  No LDS and no barriers; every derived field is computed once per
// (column, row) of the strip.  The arithmetic is the same expression for
// expression as step_kernel, so both kernels give bit-identical results.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int lane_left_i(int x) {    // value held by lane L-1 (wave_shr:1)
    return __builtin_amdgcn_update_dpp(0, x, 0x138, 0xf, 0xf, false);
}
__device__ __forceinline__ int lane_right_i(int x) {   // value held by lane L+1 (wave_shl:1)
    return __builtin_amdgcn_update_dpp(0, x, 0x130, 0xf, 0xf, false);
}
__device__ __forceinline__ double lane_left(double x) {
    const long long b = __double_as_longlong(x);
    const int lo = lane_left_i((int)(b & 0xffffffff)), hi = lane_left_i((int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
__device__ __forceinline__ double lane_right(double x) {
    const long long b = __double_as_longlong(x);
    const int lo = lane_right_i((int)(b & 0xffffffff)), hi = lane_right_i((int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
__device__ __forceinline__ float lane_left(float x) { return __int_as_float(lane_left_i(__float_as_int(x))); }
__device__ __forceinline__ float lane_right(float x) { return __int_as_float(lane_right_i(__float_as_int(x))); }

template <typename T, int NTR>
struct MRaw {            // one loaded row at this lane
    int m;
    T u, v, eta, He;
    T CH[NTR], Cb[NTR];
};

template <typename T, int NTR>
struct MRow {            // state and derived fields of one row at this lane
    int m;
    T u, v, eta, D, Du, MU, Dv, MV, Pot, plt, Ed, pll;
    T C[NTR], CH[NTR];
};

template <typename T, int VAR, int NTRC>
__global__ void __launch_bounds__(256)
march_kernel(const Dyn<T> d) {
    constexpr int NTR = NTRC > 0 ? NTRC : 1;
    const bool FN = (VAR == VAR_FN);
    const Statics<T>& a = *d.s;
    const int Nx = a.Nx, Ny = a.Ny, P = a.P;
    const int lane = threadIdx.x & 63;
    const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);   // uniform per wavefront
    const int b = xcd_tile(blockIdx.x, d.nblk);
    const int strip = b % d.ntx, grp = b / d.ntx;
    const int TY = d.ty;
    const int j0 = (grp * 4 + wid) * TY;
    if (j0 >= Ny) return;                      // uniform per wavefront
    const int jend = j0 + TY < Ny ? j0 + TY : Ny;
    const int i = strip * 62 + lane - 1;
    const unsigned gi = (unsigned)wrap(i, Nx);
    const bool out_lane = lane >= 1 && lane <= 62 && i < Nx;

    T* const st = d.state;
    const long long pl = d.plane;
    const int r0 = d.ring, r1 = d.ring == 2 ? 0 : d.ring + 1, r2 = d.ring == 0 ? 2 : d.ring - 1;
    const T* const su = st + (PL_U + d.cur) * pl;
    const T* const sv = st + (PL_V + d.cur) * pl;
    const T* const se = st + (PL_E + d.cur) * pl;

    auto load_raw = [&](int k, MRaw<T, NTR>& r) {
        const int gk = wrap(k, Ny);
        const size_t ro = (size_t)gk * P;
        r.m = ldg(a.mask + ro, gi);
        r.u = ldg(su + ro, gi);
        r.v = ldg(sv + ro, gi);
        r.eta = ldg(se + ro, gi);
        r.He = fld_r(a.He, ro, gi, gk, OE(r.m) != 0);
#pragma unroll
        for (int n = 0; n < NTRC; ++n) {
            const T* tb = st + (PL_TRC + PL_PER_TRC * n) * pl + ro;
            r.CH[n] = ldg(tb + d.cur * pl, gi);
            r.Cb[n] = ((d.cbuf >> n) & 1) ? ldg(tb + 5 * pl, gi) : T(0);
        }
    };
    auto load_hist = [&](int k, T* g, T (*gt)[2], unsigned short* lo) {
        const T* const b = st + (size_t)k * P;
        g[0] = ldg(b + (PL_GU + r0) * pl, gi); g[1] = ldg(b + (PL_GU + r1) * pl, gi);
        g[2] = ldg(b + (PL_GV + r0) * pl, gi); g[3] = ldg(b + (PL_GV + r1) * pl, gi);
        g[4] = ldg(b + (PL_GE + r0) * pl, gi); g[5] = ldg(b + (PL_GE + r1) * pl, gi);
#pragma unroll
        for (int n = 0; n < NTRC; ++n) {
            const T* tb = b + (PL_TRC + PL_PER_TRC * n + 2) * pl;
            gt[n][0] = ldg(tb + r0 * pl, gi);
            gt[n][1] = ldg(tb + r1 * pl, gi);
        }
        if constexpr (sizeof(T) == 4) {   // low parts of the FP32 state (rows of lo planes at pitch P)
            const size_t ro = (size_t)k * P;
#pragma unroll
            for (int f = 0; f < 3 + NTRC; ++f) lo[f] = 0;
            if (a.carry) {
#pragma unroll
                for (int f = 0; f < 3; ++f) lo[f] = ldg(lo_plane(st, pl, PL_LO + f) + ro, gi);
#pragma unroll
                for (int n = 0; n < NTRC; ++n)
                    lo[3 + n] = ldg(lo_plane(st, pl, PL_TRC + PL_PER_TRC * n + PL_TLO) + ro, gi);
            }
        }
    };

    MRaw<T, NTR> N;
    load_raw(j0 - 1, N);
    // One row of the march: K receives row k, P1 holds row k-1, P2 row k-2.
    auto row = [&](const int k, MRow<T, NTR>& K, MRow<T, NTR>& P1, MRow<T, NTR>& P2) {
        const int gk = wrap(k, Ny);
        // ---- row k: take the prefetched values, start the loads of row k+1
        K.m = N.m; K.u = N.u; K.v = N.v; K.eta = N.eta;
        const T He = N.He;
        T Cb[NTR];
#pragma unroll
        for (int n = 0; n < NTRC; ++n) { K.CH[n] = N.CH[n]; Cb[n] = N.Cb[n]; }
        if (k < jend) load_raw(k + 1, N);
        // AB history of row k-1, the row this call writes
        T g[6], gt[NTR][2];
        unsigned short lo[3 + NTR];
        if (k > j0) load_hist(k - 1, g, gt, lo);

        // ---- row k: fields local to the row
        K.D = FN ? (OE(K.m) ? (K.eta + He) : T(1)) : He;   // computeD; linear variants alias H_eta
#pragma unroll
        for (int n = 0; n < NTRC; ++n)   // C = CH / h, or the host-held C right after a hand-over
            K.C[n] = ((d.cbuf >> n) & 1) ? Cb[n] : dv(K.CH[n], K.D);
        const int KmL = lane_left_i(K.m);
        const T KDL = lane_left(K.D);
        K.Du = interp_noland2(K.D, KDL, OE(K.m), OE(KmL));     // eta2u_noland
        K.MU = OU(K.m) ? K.Du * K.u : T(0);                    // computeMassFluxes

        if (k >= j0) {
            const int gkm = gk == 0 ? Ny - 1 : gk - 1;
            // ---- row k: fields that need row k-1
            K.Dv = interp_noland2(K.D, P1.D, OE(K.m), OE(P1.m));   // eta2v_noland
            K.MV = OV(K.m) ? K.Dv * K.v : T(0);
            const int P1mL = lane_left_i(P1.m);
            const T P1DL = lane_left(P1.D);
            // Dh = interp4(D, eta2H_noland): (i,j), (i-1,j-1), (i-1,j), (i,j-1)
            const int m1 = OE(K.m), m2 = OE(P1mL), m3 = OE(KmL), m4 = OE(P1.m);
            const T w = noland_w<T>(m1 + m2 + m3 + m4);
            const T Dh = (((m1 ? w : T(0)) * K.D + (m2 ? w : T(0)) * P1DL) + (m3 ? w : T(0)) * KDL) +
                         (m4 ? w : T(0)) * P1.D;
            const T KvL = lane_left(K.v);
            T zeta = T(0);
            if (FN) {   // vorticityFromVelocities2D
                T za_ = T(0), zb = T(0);
                if (OV(KmL) + OV(K.m) != 0) {
                    const T z0 = ldc(a.za, 2 * (size_t)gk), z1 = ldc(a.za, 2 * (size_t)gk + 1);
                    za_ = (OH(K.m) ? z0 : z1) * (K.v - KvL);
                }
                if (OU(K.m) + OU(P1.m) != 0) {
                    const T bcH = OH(K.m) ? T(1) : a.lbc;
                    zb = dv(bcH * (ldc(a.cu, gk) * K.u - ldc(a.cu, gkm) * P1.u), a.AdT);
                }
                zeta = za_ - dv(zb, ldc(a.cH, gk));
            }
            const T pD = Dh > a.eps ? Dh : a.eps;                      // computePotVort
            K.Pot = FN ? dv(ldc(a.fH, gk) + zeta, pD) : dv(ldc(a.fH, gk), pD);
            T plt = T(0);                                              // P_lambda-theta on ocean H
            if (a.mix && OH(K.m)) {
                const T* c = a.pltc + (size_t)gk * 12;
                const unsigned q = (OU(K.m) * 2 + OU(P1.m)) * 3;
                plt = Dh * ((ldg(c, q) * (K.v - KvL) + ldg(c, q + 1) * K.u) + ldg(c, q + 2) * P1.u);
            }
            K.plt = plt;
            // ---- row k-1: fields that need row k
            const T P1uR = lane_right(P1.u);
            T Ed = T(0);                                               // computeEdens
            if (OE(P1.m)) {
                if (FN) {
                    Ed = ((T(0.5) * (P1uR * P1uR) + T(0.5) * (P1.u * P1.u)) +
                          (T(0.5) * (K.v * K.v) + T(0.5) * (P1.v * P1.v))) / T(2) + a.G * P1.eta;
                } else {
                    Ed = a.G * P1.eta;
                }
            }
            P1.Ed = Ed;
            T pll = T(0);                                              // P_lambda-lambda on ocean eta
            if (a.mix && OE(P1.m)) {
                const T* c = a.pllc + (size_t)gkm * 12;
                const unsigned q = (OV(K.m) * 2 + OV(P1.m)) * 3;
                pll = P1.D * ((ldg(c, q) * (P1uR - P1.u) + ldg(c, q + 1) * K.v) + ldg(c, q + 2) * P1.v);
            }
            P1.pll = pll;

            // ---- outputs of row j = k-1 from rows j-1 (P2), j (P1), j+1 (K)
            const int j = k - 1;
            if (j >= j0) {
                // neighbour values, taken with every lane active
                const int P1mR = lane_right_i(P1.m);
                const T P1MUR = lane_right(P1.MU), P2MUR = lane_right(P2.MU);
                const T P1MVL = lane_left(P1.MV), KMVL = lane_left(K.MV);
                const T P1EdL = lane_left(P1.Ed), P1pllL = lane_left(P1.pll);
                const T P1PotR = lane_right(P1.Pot), P1pltR = lane_right(P1.plt);
                T P1vL = T(0), P2uR = T(0);
                if (a.quad) { P1vL = lane_left(P1.v); P2uR = lane_right(P2.u); }
                T P1DR = T(0), CHe[NTR], CHw[NTR], Ce[NTR], Cw[NTR];
                if (NTRC > 0) P1DR = lane_right(P1.D);
#pragma unroll
                for (int n = 0; n < NTRC; ++n) {
                    CHe[n] = lane_right(P1.CH[n]); CHw[n] = lane_left(P1.CH[n]);
                    Ce[n] = lane_right(P1.C[n]); Cw[n] = lane_left(P1.C[n]);
                }
                if (out_lane) {
                    const size_t ro = (size_t)j * P;
                    const int jp = gk;   // j+1, wrapped
                    const int m = P1.m;
                    const int oe = OE(m), ou = OU(m), ov = OV(m), oh = OH(m);
                    const T Du = P1.Du, Dv = P1.Dv;
                    // SWM_forcing_update
                    T Fe = fld_r(a.Fec, ro, gi, j, oe);
                    if (a.has_oe) Fe = Fe + (T(2) * (fld_r(a.oRe, ro, gi, j, oe) * d.oc[2] - fld_r(a.oIe, ro, gi, j, oe) * d.os[2])) * T(oe);
                    T Fx = fld_r(a.Fxc, ro, gi, j, ou);
                    if (a.has_taux && ou) {
                        const T tx = fld_r(a.taux, ro, gi, j, true);
                        Fx = Fx + dv(a.has_tau_scale ? a.tau_scale * tx : tx, a.RHO0 * Du);
                    }
                    if (a.has_ox) Fx = Fx + (T(2) * (fld_r(a.oRx, ro, gi, j, ou) * d.oc[0] - fld_r(a.oIx, ro, gi, j, ou) * d.os[0])) * T(ou);
                    T Fy = fld_r(a.Fyc, ro, gi, j, ov);
                    if (a.has_tauy && ov) {
                        const T ty = fld_r(a.tauy, ro, gi, j, true);
                        Fy = Fy + dv(a.has_tau_scale ? a.tau_scale * ty : ty, a.RHO0 * Dv);
                    }
                    if (a.has_oy) Fy = Fy + (T(2) * (fld_r(a.oRy, ro, gi, j, ov) * d.oc[1] - fld_r(a.oIy, ro, gi, j, ov) * d.os[1])) * T(ov);
                    if (a.nstream) apply_streams<T>(a, d, ro + gi, ou, ov, oe, Du, Dv, Fx, Fy, Fe);

                    // SWM_timestep_nonlinear
                    T Ge = T(0), Gu = T(0), Gv = T(0);
                    if (oe) {
                        const T t1 = dv(-(P1MUR - P1.MU), a.dL);
                        const T t2 = dv(K.MV * ldc(a.cv, jp) - P1.MV * ldc(a.cv, j), a.dT);
                        Ge = dv(t1 - t2, ldc(a.ace, j)) + Fe;
                    }
                    if (ou) {
                        const T P2u = T(0.5) * K.Pot + T(0.5) * P1.Pot;
                        const T M4 = ((T(0.25) * K.MV + T(0.25) * P1MVL) + T(0.25) * KMVL) + T(0.25) * P1.MV;
                        T gg = P2u * M4;
                        gg = gg - dv(P1.Ed - P1EdL, ldc(a.edu, j));
                        if (a.quad) {
                            const T vi = ((T(0.25) * K.v + T(0.25) * P1vL) + T(0.25) * KvL) + T(0.25) * P1.v;
                            T qf = fld_r(a.gsu, ro, gi, j, true) * sqrt(P1.u * P1.u + (vi * vi)) * P1.u;
                            if (a.barotropic) qf = dv(qf, fld_r(a.Hu, ro, gi, j, true));
                            gg = gg - qf;
                        }
                        if (a.mix) {
                            const int ot = (oh + OH(K.m)) < 1 ? 1 : (oh + OH(K.m));
                            const T* c = a.lmu + (size_t)j * 6;
                            const T c0 = ot == 1 ? ldc(c, 0) : ldc(c, 3), c1 = ot == 1 ? ldc(c, 1) : ldc(c, 4),
                                    c2 = ot == 1 ? ldc(c, 2) : ldc(c, 5);
                            gg = gg + dv((c0 * (P1.pll - P1pllL) + c1 * K.plt) + c2 * P1.plt, Du);
                        }
                        Gu = gg + Fx;
                    }
                    if (ov) {
                        const T P2v = T(0.5) * P1PotR + T(0.5) * P1.Pot;
                        const T M4 = ((T(0.25) * P1MUR + T(0.25) * P2.MU) + T(0.25) * P1.MU) + T(0.25) * P2MUR;
                        T gg = -(P2v * M4);
                        gg = gg - dv(P1.Ed - P2.Ed, a.AdT);
                        if (a.quad) {
                            const T ui = ((T(0.25) * P1uR + T(0.25) * P2.u) + T(0.25) * P1.u) + T(0.25) * P2uR;
                            T qf = fld_r(a.gsv, ro, gi, j, true) * sqrt(P1.v * P1.v + (ui * ui)) * P1.v;
                            if (a.barotropic) qf = dv(qf, fld_r(a.Hv, ro, gi, j, true));
                            gg = gg - qf;
                        }
                        if (a.mix) {
                            const int ot = (oe + OE(P2.m)) < 1 ? 1 : (oe + OE(P2.m));
                            const T* c = a.lmv + (size_t)j * 6;
                            const T c0 = ot == 1 ? ldc(c, 0) : ldc(c, 3), c1 = ot == 1 ? ldc(c, 1) : ldc(c, 4),
                                    c2 = ot == 1 ? ldc(c, 2) : ldc(c, 5);
                            gg = gg + dv((c0 * (P1pltR - P1.plt) + c1 * P1.pll) + c2 * P2.pll, Dv);
                        }
                        Gv = gg + Fy;
                    }
                    // integrate_AB; land keeps the previous newest tendency
                    const T Gu3 = ou ? Gu : g[1];
                    const T Gv3 = ov ? Gv : g[3];
                    const T Ge3 = oe ? Ge : g[5];
                    const T dvu = fld_r(a.imu, ro, gi, j, ou), dvv = fld_r(a.imv, ro, gi, j, ov),
                            dve = fld_r(a.ime, ro, gi, j, oe);
                    T un, vn, en;
                    unsigned short lu = 0, lv = 0, le = 0;
                    if constexpr (sizeof(T) == 4) {   // increments added to the carried pair
                        T iu = d.dtc1 * g[0], iv = d.dtc1 * g[2], ie = d.dtc1 * g[4];
                        if (d.tstep >= 2) {
                            iu = madd(iu, d.dtc2, g[1]);
                            iv = madd(iv, d.dtc2, g[3]);
                            ie = madd(ie, d.dtc2, g[5]);
                        }
                        if (d.tstep >= 3) {
                            iu = madd(iu, d.dtc3, Gu3);
                            iv = madd(iv, d.dtc3, Gv3);
                            ie = madd(ie, d.dtc3, Ge3);
                        }
                        un = carry_step(P1.u, lo[0], iu, dvu, lu);   // FP32: dvu etc. hold 1 - 1/impl
                        vn = carry_step(P1.v, lo[1], iv, dvv, lv);
                        en = carry_step(P1.eta, lo[2], ie, dve, le);
                    } else {
                        un = P1.u, vn = P1.v, en = P1.eta;
                        un = madd(un, d.dtc1, g[0]);
                        vn = madd(vn, d.dtc1, g[2]);
                        en = madd(en, d.dtc1, g[4]);
                        if (d.tstep >= 2) {
                            un = madd(un, d.dtc2, g[1]);
                            vn = madd(vn, d.dtc2, g[3]);
                            en = madd(en, d.dtc2, g[5]);
                        }
                        if (d.tstep >= 3) {
                            un = madd(un, d.dtc3, Gu3);
                            vn = madd(vn, d.dtc3, Gv3);
                            en = madd(en, d.dtc3, Ge3);
                        }
                        un = dv(un, dvu);
                        vn = dv(vn, dvv);
                        en = dv(en, dve);
                    }
                    if (FN) {   // where (eta < minD - H) eta = minD - H
                        const T lim = a.minD - fld_r(a.He, ro, gi, j, oe);
                        if (en < lim) { en = lim; le = 0; }
                    }
                    T* const wb = st + ro;
                    stg(wb + (PL_U + 1 - d.cur) * pl, gi, un);
                    stg(wb + (PL_V + 1 - d.cur) * pl, gi, vn);
                    stg(wb + (PL_E + 1 - d.cur) * pl, gi, en);
                    if constexpr (sizeof(T) == 4) {
                        if (a.carry) {
                            stg((unsigned short*)lo_plane(st, pl, PL_LO) + ro, gi, lu);
                            stg((unsigned short*)lo_plane(st, pl, PL_LO + 1) + ro, gi, lv);
                            stg((unsigned short*)lo_plane(st, pl, PL_LO + 2) + ro, gi, le);
                        }
                    }
                    stg(wb + (PL_GU + r2) * pl, gi, Gu3);
                    stg(wb + (PL_GV + r2) * pl, gi, Gv3);
                    stg(wb + (PL_GE + r2) * pl, gi, Ge3);

                    // TRC_tracer_incr + TRC_tracer_integrate
#pragma unroll
                    for (int n = 0; n < NTRC; ++n) {
                        T* const tb = st + (PL_TRC + PL_PER_TRC * n) * pl + ro;
                        const T CH = P1.CH[n], CHn = K.CH[n], CHs = P2.CH[n];
                        const T C = P1.C[n], Cn = K.C[n], Cs = P2.C[n];
                        // tracer velocities are the master fields u = 0 + SWM_u (model_advance)
                        const T um = T(0) + P1.u, ume = T(0) + P1uR, vm = T(0) + P1.v, vmn = T(0) + K.v;
                        const int oee = OE(P1mR), oew = OE(P1mL), oen = OE(K.m), oes = OE(P2.m);
                        const int oue = OU(P1mR), ovn = OV(K.m);
                        T diff = T(0);
                        if (oe) {
                            const T kap = a.kappa[n];
                            const T hue = interp_noland2(P1DR, P1.D, oee, oe);
                            const T hvn = K.Dv;   // interp of D(i,j+1), D(i,j) with the same weights
                            const T c9 = oue ? ldc(a.t9, j) : T(0.0);
                            const T c10 = ou ? -ldc(a.t9, j) : T(-0.0);
                            const T c11 = ovn ? ldc(a.t11, j) : T(0.0);
                            const T c12 = ov ? -ldc(a.t12, j) : T(-0.0);
                            diff = ((((kap * hue) * (Ce[n] - C)) * c9 + ((kap * Du) * (C - Cw[n])) * c10) +
                                    ((kap * hvn) * (Cn - C)) * c11) + ((kap * Dv) * (C - Cs)) * c12;
                        }
                        const T forcing = C * Fe;
                        const T c1 = trc_adv_r<T>(true, oue & oee, oee + oe, a.tA, j);
                        const T c2 = trc_adv_r<T>(false, ou & oew, oew + oe, a.tA, j);
                        const T c3 = trc_adv_r<T>(true, ovn & oen, oen + oe, a.tB, j);
                        const T c4 = trc_adv_r<T>(false, ov & oes, oes + oe, a.tC, j);
                        const T c5 = trc_adv_r<T>(true, oue & oe, oee + oe, a.tA, j);
                        const T c6 = trc_adv_r<T>(false, ou & oe, oew + oe, a.tA, j);
                        const T c7 = trc_adv_r<T>(true, ovn & oe, oen + oe, a.tB, j);
                        const T c8 = trc_adv_r<T>(false, ov & oe, oes + oe, a.tC, j);
                        T gg = (CHe[n] * ume) * c1 + (CHw[n] * um) * c2;
                        gg = gg + (CHn * vmn) * c3;
                        gg = gg + (CHs * vm) * c4;
                        gg = gg + (CH * ume) * c5;
                        gg = gg + (CH * um) * c6;
                        gg = gg + (CH * vmn) * c7;
                        gg = gg + (CH * vm) * c8;
                        gg = gg + diff;
                        gg = gg - (fld_r(a.gC[n], ro, gi, j, oe) * P1.D) * (C - fld_r(a.C0[n], ro, gi, j, oe));
                        gg = gg + forcing;
                        const T dvc = fld_r(a.timpl[n], ro, gi, j, oe);
                        T chn;
                        if constexpr (sizeof(T) == 4) {
                            T inc = d.dtc1 * gt[n][0];
                            if (d.tstep >= 2) inc = madd(inc, d.dtc2, gt[n][1]);
                            if (d.tstep >= 3) inc = madd(inc, d.dtc3, gg);
                            unsigned short lc;
                            chn = carry_step(CH, lo[3 + n], inc, dvc, lc);
                            if (a.carry)
                                stg((unsigned short*)lo_plane(st, pl, PL_TRC + PL_PER_TRC * n + PL_TLO) + ro, gi, lc);
                        } else {
                            chn = CH + d.dtc1 * gt[n][0];
                            if (d.tstep >= 2) chn = madd(chn, d.dtc2, gt[n][1]);
                            if (d.tstep >= 3) chn = madd(chn, d.dtc3, gg);
                            chn = dv(chn, dvc);
                        }
                        stg(tb + (1 - d.cur) * pl, gi, chn);
                        stg(tb + (2 + r2) * pl, gi, gg);
                    }
                }
            }
        }
    };
    // rows j0-1 .. jend, unrolled by three so the roles rotate without copies
    MRow<T, NTR> A, B, C;
    for (int k = j0 - 1; k <= jend; k += 3) {
        row(k, A, C, B);
        if (k + 1 <= jend) row(k + 1, B, A, C);
        if (k + 2 <= jend) row(k + 2, C, B, A);
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
    const T* c5r;               // 1 / c5t, correctly rounded (dvr)
    double* partial;            // per-block partial sums of res^2 (or r^2)
    int* done;                  // set once converged: later sweeps return at once
    int residual_mode;          // 0: sum res^2 (reference), 1: sum (L chi - B)^2
    int accumulate;             // 0: no convergence test after this sweep (no partial sums)
    // blocked sweeps: block 0 runs the previous launch's convergence test
    const double* prev_partial; // that launch's partial sums (gridDim.x of them)
    int prev_check, prev_sweep; // whether it had a test, and the sweep tested
    double eps, bnorm;
    long long* iters;
    double* norm_out;
    // reference mode with K sweeps per launch (mid): the chi of the output
    // cells after each of the first K - 1 sweeps and their residual partials
    // are kept as well, and the next launch's block 0 tests the sweeps in order
    int mid, prev_mid;
    T* chi_mid[3];
    double* partial_mid[3];
    const double* prev_partial_mid[3];
};

// The convergence test of one sweep, by the 256 threads of one block: sum the
// n partials in a fixed order, take the norm (mode 0: sqrt(sum), the
// reference's; mode 1: sqrt(sum) / bnorm), record it and the sweep, and set
// *done when it is below eps.  Returns that outcome to every thread.
__device__ __forceinline__ bool sor_test(const double* partial, int n, int sweep, double eps, double bnorm, int mode,
                                         int* done, long long* iters, double* norm_out, double* red, int* flag) {
    double acc = 0.0;
    for (int b = threadIdx.x; b < n; b += 256) acc += partial[b];
    for (int s = 32; s >= 1; s >>= 1) acc += __shfl_xor(acc, s, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        const double tot = ((red[0] + red[1]) + red[2]) + red[3];
        double nrm = sqrt(tot);
        if (mode == 1) nrm = bnorm > 0.0 ? nrm / bnorm : nrm;
        *norm_out = nrm;
        *iters = sweep;
        *flag = nrm < eps;
        if (nrm < eps) *done = 1;
    }
    __syncthreads();
    return *flag != 0;
}

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
                const int c5i = ((j * 3 + ouip + ou) * 2 + ovjp) * 2 + ov;
                const T c5 = a.c5t[c5i];
                T* chi = a.chi;
                const T x = chi[off];
                const T lhs = (((c1 * chi[(size_t)j * a.P + ip] + c2 * chi[(size_t)j * a.P + im]) +
                                c3 * chi[(size_t)jp * a.P + i]) + c4 * chi[(size_t)jm * a.P + i]) + c5 * x;
                const T r = lhs - a.B[off];
                T res;
                if (a.relax) {
                    int h = 2 * (a.sweep - 1) + a.color;
                    if (h >= a.nrelax) h = a.nrelax - 1;
                    res = dvr(T(a.relax[h]) * r, c5, a.c5r[c5i]);
                } else {
                    res = dvr(T(a.omega) * r, c5, a.c5r[c5i]);
                }
                chi[off] = x - res;
                const double q = a.residual_mode ? (double)r : (double)res;
                acc = q * q;
            }
        }
    }
    if (!a.accumulate) return;   // uniform: no barrier is skipped by part of the block
    // fixed-order block reduction: wavefront butterfly then the 4 wave sums in order
    for (int s = 32; s >= 1; s >>= 1) acc += __shfl_xor(acc, s, 64);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0) red[wv] = acc;
    __syncthreads();
    if (threadIdx.x == 0) a.partial[blockIdx.x] = ((red[0] + red[1]) + red[2]) + red[3];
}

// K sweeps in one launch (fixed omega, or the reference's Chebyshev omega per
// half sweep) on a TX x TY tile held in LDS with a 2K-cell halo (read from
// chi_in, written to chi_out: blocks read their halos while neighbours write).
// Half sweep hs updates the colour-(hs % 2) cells at least hs + 1 cells inside
// the staged block, whose dependency cones stay inside it, so the K sweeps give
// each output cell exactly the values of K global red-black sweeps (sor_sweep,
// same arithmetic).  With accumulate the last sweep's residuals of the output
// cells (r, or the reference's res) are summed in FP64 in a fixed order into
// partial[block]; block 0 first runs the previous launch's convergence test.
//
// A cell update reads LDS only: the per-row coefficients of the staged rows
// are staged with the block as a table in which each of c1, c3, c4 sits beside
// a +0 and is picked by its mask bit as an index (the value sor_sweep's select
// gives, without a branch around the read).  Where the colours alternate
// regularly a half sweep walks exactly the cells of its colour inside its
// shrinking region, and each thread forms the new values of all its cells
// before it stores any of them, so their LDS reads overlap.
template <typename T, int TX, int TY, int K>
__global__ void __launch_bounds__(256) sor_block(const SorArgs<T> a, const T* __restrict__ chi_in,
                                                 T* __restrict__ chi_out, int ntx) {
    constexpr int H = 2 * K, W = TX + 2 * H, HH = TY + 2 * H, E = W * HH;
    // per staged row: {0, c1}, {0, c3}, {0, c4}, c5[12], 1/c5[12]
    constexpr int TC1 = 0, TC3 = 2, TC4 = 4, TC5 = 6, TC5R = 18, NC = 30, NL = 27;
    static_assert(NL == 3 + 2 * 12 && NC == TC5R + 12, "row table layout");
    __shared__ T s_chi[E];
    __shared__ T s_B[E];
    __shared__ T s_c[HH][NC];
    // per cell: ov | ov(i,j+1) << 1 | (ou(i+1,j) + ou) << 2 | ou(i+1,j) << 4 | ou << 5 |
    // ocean eta << 6 | colour ((i + j) % 2) << 7: the low four bits index c5
    __shared__ uint8_t s_m[E];
    __shared__ double red[256 / 64];
    const int Nx = a.Nx, Ny = a.Ny, P = a.P;
    const int i0 = (blockIdx.x % ntx) * TX, j0 = (blockIdx.x / ntx) * TY;
    const int tid = threadIdx.x;
    // Block 0 tests the previous launch's sweep while the others sweep on; once
    // a test has passed, later launches return at once and the converged chi
    // stays in the buffer that launch wrote (the host picks it by the sweep).
    __shared__ int conv;
    if (a.prev_check && blockIdx.x == 0 && !*a.done) {
        const int mode = a.residual_mode ? 1 : 0;
        if (a.prev_mid)
            for (int sw = 0; sw < K - 1; ++sw)
                if (sor_test(a.prev_partial_mid[sw], (int)gridDim.x, a.prev_sweep - (K - 1) + sw, a.eps, a.bnorm, mode,
                             a.done, a.iters, a.norm_out, red, &conv))
                    return;
        if (sor_test(a.prev_partial, (int)gridDim.x, a.prev_sweep, a.eps, a.bnorm, mode, a.done, a.iters,
                     a.norm_out, red, &conv))
            return;
    }
    if (*a.done) return;
    // every global load of the stage is issued before the first LDS store, so
    // the block waits for one memory latency, not one per cell it stages
    constexpr int NS = (E + 255) / 256, NQ = (HH * NL + 255) / 256;
    T rchi[NS], rB[NS], rc[NQ];
    int rm[NS];
#pragma unroll
    for (int q = 0; q < NS; ++q) {
        const int e = tid + q * 256;
        if (e < E) {
            const int lx = e % W, ly = e / W;
            const int gi = wrap(i0 - H + lx, Nx), gj = wrap(j0 - H + ly, Ny);
            const size_t off = (size_t)gj * P + gi;
            const int gip = gi + 1 == Nx ? 0 : gi + 1, gjp = gj + 1 == Ny ? 0 : gj + 1;
            rchi[q] = chi_in[off];
            rB[q] = a.B[off];
            const int mc = a.mask[off], mip = a.mask[(size_t)gj * P + gip], mjp = a.mask[(size_t)gjp * P + gi];
            const int ouip = (mip >> 1) & 1, ou = (mc >> 1) & 1, ovjp = (mjp >> 2) & 1, ov = (mc >> 2) & 1;
            rm[q] = ov | ovjp << 1 | (ouip + ou) << 2 | ouip << 4 | ou << 5 | ((mc & OC_E) ? 1 << 6 : 0) |
                    ((gi + gj) & 1) << 7;
        }
    }
    // the 3 + 12 + 12 values of each staged row
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        const int x = tid + q * 256;
        if (x < HH * NL) {
            const int ly = x / NL, k = x % NL;
            const size_t gj = (size_t)wrap(j0 - H + ly, Ny);
            rc[q] = k == 0 ? a.c1t[gj] : k == 1 ? a.c3t[gj] : k == 2 ? a.c4t[gj]
                  : k < 15 ? a.c5t[gj * 12 + (k - 3)] : a.c5r[gj * 12 + (k - 15)];
        }
    }
#pragma unroll
    for (int q = 0; q < NS; ++q) {
        const int e = tid + q * 256;
        if (e < E) {
            s_chi[e] = rchi[q];
            s_B[e] = rB[q];
            s_m[e] = (uint8_t)rm[q];
        }
    }
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        const int x = tid + q * 256;
        if (x < HH * NL) {
            const int ly = x / NL, k = x % NL;
            s_c[ly][k < 3 ? 2 * k + 1 : k + 3] = rc[q];
            if (k < 3) s_c[ly][2 * k] = T(0);
        }
    }
    __syncthreads();
    // Without an odd periodic wrap inside the block the colours alternate
    // along every staged row (colour = (p0 + lx + ly) % 2): each half-sweep
    // then visits only the cells of its colour; otherwise it walks the block
    // and tests the colour bit.
    static_assert(W % 2 == 0, "pairs of cells along a row");
    const bool regular = ((Nx & 1) == 0 || (i0 - H >= 0 && i0 - H + W <= Nx)) &&
                         ((Ny & 1) == 0 || (j0 - H >= 0 && j0 - H + HH <= Ny));
    const int p0 = (i0 + j0) & 1;
    double acc = 0.0, acc_mid[3] = {0.0, 0.0, 0.0};
    // the update of cell (lx, ly) (sor_sweep's arithmetic): its new value, and
    // the residual term r or res; land cells keep their value
    auto upd = [&](int lx, int ly, T om, T& nv, T& rq) -> bool {
        const int e = ly * W + lx;
        const int m = s_m[e];
        const T* const c = s_c[ly];
        const T c1 = c[TC1 + ((m >> 4) & 1)], c2 = c[TC1 + ((m >> 5) & 1)], c3 = c[TC3 + ((m >> 1) & 1)],
                c4 = c[TC4 + (m & 1)];
        const int c5i = m & 15;   // ((ou(i+1,j) + ou) * 2 + ov(i,j+1)) * 2 + ov
        const T c5 = c[TC5 + c5i];
        const T x = s_chi[e];
        const T lhs = (((c1 * s_chi[e + 1] + c2 * s_chi[e - 1]) + c3 * s_chi[e + W]) + c4 * s_chi[e - W]) + c5 * x;
        const T r = lhs - s_B[e];
        const T res = dvr(om * r, c5, c[TC5R + c5i]);
        const bool oc = (m >> 6) & 1;
        nv = oc ? x - res : x;
        rq = a.residual_mode ? r : res;
        return oc;
    };
    // tested half sweeps sum the residuals of the output cells
    auto add = [&](double& sum, int lx, int ly, T rq) {
        if (lx >= H && lx < H + TX && ly >= H && ly < H + TY && i0 + lx - H < Nx && j0 + ly - H < Ny) {
            const double q = (double)rq;
            sum += q * q;
        }
    };
    auto store_out = [&](T* out) {   // the output cells of the tile
        for (int o = tid; o < TX * TY; o += 256) {
            const int i = i0 + o % TX, j = j0 + o / TX;
            if (i < Nx && j < Ny) out[(size_t)j * P + i] = s_chi[(H + o / TX) * W + H + o % TX];
        }
    };
    constexpr int NIM = (HH * (W / 2) + 255) / 256;   // cells of one colour per thread, at most
#pragma unroll
    for (int hs = 0; hs < 2 * K; ++hs) {
        const int color = hs & 1;
        const int lo = hs + 1;
        const bool tail = a.accumulate && hs >= 2 * K - 2;
        const int sw = hs / 2;                          // the sweep of this half sweep
        const bool inner = a.mid && sw < K - 1;         // an earlier sweep, tested too (mid)
        // omega of this half-sweep: the reference's Chebyshev schedule (held
        // once stationary) or the fixed omega; a.sweep is the launch's first sweep
        T om = T(a.omega);
        if (a.relax) {
            int h = 2 * (a.sweep - 1) + hs;
            if (h >= a.nrelax) h = a.nrelax - 1;
            om = T(a.relax[h]);
        }
        if (regular) {
            // rows lo .. HH-1-lo, (W - 2 lo) / 2 cells of the colour in each
            const int RW2 = (W - 2 * lo) / 2, n = (HH - 2 * lo) * RW2;
            T nv[NIM], rq[NIM];
            int lxs[NIM], lys[NIM];
            bool oc[NIM];
            // slots past the region's cells (k * 256 >= n, known once the
            // half sweeps are unrolled) are dropped; in the last one, threads
            // past the last cell repeat it and drop the result, so every
            // slot's reads are issued unconditionally
#pragma unroll
            for (int k = 0; k < NIM; ++k) {
                oc[k] = false;
                if (k * 256 >= n) continue;
                const int q = tid + k * 256 < n ? tid + k * 256 : n - 1;
                lys[k] = lo + q / RW2;
                lxs[k] = lo + 2 * (q % RW2) + ((color + p0 + lys[k] + lo) & 1);
                oc[k] = upd(lxs[k], lys[k], om, nv[k], rq[k]) && tid + k * 256 < n;
            }
#pragma unroll
            for (int k = 0; k < NIM; ++k) {
                if (k * 256 < n && oc[k]) {
                    s_chi[lys[k] * W + lxs[k]] = nv[k];
                    if (tail) add(acc, lxs[k], lys[k], rq[k]);
                    if (inner) add(acc_mid[sw < 3 ? sw : 0], lxs[k], lys[k], rq[k]);
                }
            }
        } else {
            for (int q = tid; q < E; q += 256) {
                const int lx = q % W, ly = q / W;
                if (lx < lo || lx > W - 1 - lo || ly < lo || ly > HH - 1 - lo) continue;
                if ((s_m[q] >> 7) != color) continue;
                T nv, rq;
                if (upd(lx, ly, om, nv, rq)) {
                    s_chi[q] = nv;
                    if (tail) add(acc, lx, ly, rq);
                    if (inner) add(acc_mid[sw < 3 ? sw : 0], lx, ly, rq);
                }
            }
        }
        __syncthreads();
        if ((hs & 1) && inner) {   // chi after sweep sw, before the next sweep changes it
            store_out(a.chi_mid[sw < 3 ? sw : 0]);
            __syncthreads();
        }
    }
    store_out(chi_out);
    if (!a.accumulate) return;   // uniform
    // fixed-order block sum: wavefront butterfly, then the four wave sums in order
    auto block_sum = [&](double v) -> double {
        for (int s = 32; s >= 1; s >>= 1) v += __shfl_xor(v, s, 64);
        if ((tid & 63) == 0) red[tid >> 6] = v;
        __syncthreads();
        const double t = ((red[0] + red[1]) + red[2]) + red[3];
        __syncthreads();
        return t;
    };
    const double tot = block_sum(acc);
    if (tid == 0) a.partial[blockIdx.x] = tot;
    if (a.mid) {
#pragma unroll
        for (int q = 0; q < K - 1 && q < 3; ++q) {
            const double tot_mid = block_sum(acc_mid[q]);
            if (tid == 0) a.partial_mid[q][blockIdx.x] = tot_mid;
        }
    }
}

// Sum the block partials in a fixed order (one block), test convergence.
// mode 0: sqrt(sum) < eps (reference); mode 1: sqrt(sum)/bnorm < eps.
__global__ void __launch_bounds__(256)
sor_check(const double* partial, int nblk, int sweep, double eps, double bnorm, int mode,
          int* done, long long* iters, double* norm_out) {
    __shared__ double red[4];
    __shared__ int conv;
    if (*done) return;
    sor_test(partial, nblk, sweep, eps, bnorm, mode, done, iters, norm_out, red, &conv);
}

// computeDivergence(u, v, div) (src/calc_lib.f90:607-622), evaluated in the
// field precision T and stored in TO (the SOR right-hand side precision).
template <typename T, typename TO = T>
__global__ void __launch_bounds__(256)
divergence_kernel(int Nx, int Ny, int P, const T* u, const T* v, const uint8_t* mask, const T* dza,
                  const T* cv, const T* ce, T AdT, T lbc, TO* div) {
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
    div[off] = (TO)(za + zb / ce[j]);
}

// A master field after model_advance (src/model.f90:238-259): 0, plus the
// grid's ocean mask times the DYNFROMFILE dataset (DFF_advance,
// src/dynFromFile_module.f90:137-160), plus the shallow-water field at N0.
template <typename T>
__global__ void __launch_bounds__(256)
master_kernel(int Nx, int Ny, int P, const uint8_t* mask, uint8_t bit, const T* rec, SStep<T> st, const T* swm,
              T* out) {
    const long idx = (long)blockIdx.x * 256 + threadIdx.x;
    if (idx >= (long)Nx * Ny) return;
    const int j = (int)(idx / Nx), i = (int)(idx - (long)j * Nx);
    const size_t off = (size_t)j * P + i;
    const T o = (mask[off] & bit) ? T(1) : T(0);
    out[off] = (T(0) + o * dyn_val(rec, st, off)) + swm[off];
}

// sum over ocean eta points of B^2 in FP64, fixed order: per-block partials
// (one 256-thread tree per block) then one block over the partials.
template <typename TS>
__global__ void __launch_bounds__(256)
sq_norm_kernel(int Nx, int Ny, int P, const TS* B, const uint8_t* mask, double* partial) {
    __shared__ double red[256];
    double acc = 0.0;
    const long n = (long)Nx * Ny;
    for (long idx = (long)blockIdx.x * 256 + threadIdx.x; idx < n; idx += (long)gridDim.x * 256) {
        const int j = (int)(idx / Nx), i = (int)(idx - (long)j * Nx);
        const size_t off = (size_t)j * P + i;
        if (mask[off] & OC_E) {
            const double b = (double)B[off];
            acc += b * b;
        }
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s2 = 128; s2 > 0; s2 >>= 1) {
        if ((int)threadIdx.x < s2) red[threadIdx.x] += red[threadIdx.x + s2];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

__global__ void __launch_bounds__(256)
sq_norm_final(const double* partial, int nblk, double* out) {
    __shared__ double red[256];
    double acc = 0.0;
    for (int b = threadIdx.x; b < nblk; b += 256) acc += partial[b];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s2 = 128; s2 > 0; s2 >>= 1) {
        if ((int)threadIdx.x < s2) red[threadIdx.x] += red[threadIdx.x + s2];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = sqrt(red[0]);
}

// ---------------------------------------------------------------------------
// Averaging buffer of a diag task (addDataToTaskBuffer, src/diagTask.f90:318-340):
// buf = buf + w * f**p, p = 1 or 2, in FP64 whatever the field precision.
// ---------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(256)
avg_accumulate(int Nx, int Ny, int P, const T* f, double* buf, double w, int square) {
    const long idx = (long)blockIdx.x * 256 + threadIdx.x;
    if (idx >= (long)Nx * Ny) return;
    const int j = (int)(idx / Nx), i = (int)(idx - (long)j * Nx);
    const size_t off = (size_t)j * P + i;
    const double x = (double)f[off];
    buf[off] = buf[off] + w * (square ? x * x : x);
}

// ---------------------------------------------------------------------------
// Conservation sums in FP64 with compensated (Neumaier) accumulation, for the
// FP32 storage path as well as the FP64 one.  Over ocean eta points, weighted
// by cos(lat_eta):
//   [0] volume  sum eta
//   [1] energy  sum 0.5*D*(0.5*(u(i)^2 + u(i+1)^2) + 0.5*(v(j)^2 + v(j+1)^2)) + 0.5*G*eta^2
//   [2] tracer  sum CH
// Cells are visited in a fixed order per thread and combined in a fixed tree,
// so the result is reproducible run to run for a given grid.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void nsum_add(double& s, double& c, double x) {
    const double t = s + x;
    if (fabs(s) >= fabs(x)) c += (s - t) + x;
    else c += (x - t) + s;
    s = t;
}

// computeStreamfunction (src/calc_lib.f90:727-780) of the master u, v at N0
// (without an elliptic solver u_nd = u, 563-564).  Path integrals in
// FP64, each SUM from its first element to its last as the reference's loop
// runs it.  psi_prep writes the summands: pv = [H_v*]v, with the ocean_v
// factor in row 1 (H_v*ocean_v*v); pu = [H_u*]u, with ocean_u in column Nx.
// computeGradient (src/calc_lib.f90:638-648) of the velocity potential on the
// eta grid: pder_zonal onto the u grid with the points (i, i-1) and
// pder_meridional onto the v grid with (j, j-1) (pder_zonal3D and
// pder_meridional3D, 892-998; getOutGrid, 1220-1310), zero where both eta
// points are land, scaled by the out grid's bc (1 on its ocean points, lbc on
// its land points).  gza[2j] = 1 / (A cos_u(j) dLambda), gza[2j + 1] =
// lbc / (A cos_u(j) dLambda), the reference's bc / (A*cos*dLambda).  Then
// computeNonDivergentFlowField (548-590): u_nd = u - du, v_nd = v - dv, on
// every point, into und / vnd (FP64).
template <typename T, typename TS>
__global__ void __launch_bounds__(256)
nondivergent_kernel(int Nx, int Ny, int P, const uint8_t* mask, const T* u, const T* v, const TS* chi,
                    const double* gza, double lbc, double AdT, double* und, double* vnd) {
    const long idx = (long)blockIdx.x * 256 + threadIdx.x;
    if (idx >= (long)Nx * Ny) return;
    const int j = (int)(idx / Nx), i = (int)(idx - (long)j * Nx);
    const int im = i == 0 ? Nx - 1 : i - 1, jm = j == 0 ? Ny - 1 : j - 1;
    const size_t off = (size_t)j * P + i;
    const uint8_t m = mask[off], mim = mask[(size_t)j * P + im], mjm = mask[(size_t)jm * P + i];
    double du = 0.0, dv_ = 0.0;
    if ((mim & OC_E) || (m & OC_E)) du = gza[2 * j + ((m & OC_U) ? 0 : 1)] * ((double)chi[off] - (double)chi[(size_t)j * P + im]);
    if ((m & OC_E) || (mjm & OC_E)) {
        const double bc = (m & OC_V) ? 1.0 : lbc;
        dv_ = bc * ((double)chi[off] - (double)chi[(size_t)jm * P + i]) / AdT;
    }
    und[off] = (double)u[off] - du;
    vnd[off] = (double)v[off] - dv_;
}

// und / vnd (optional): the non-divergent flow of the elliptic-solver build
// (computeStreamfunction calls computeNonDivergentFlowField first,
// src/calc_lib.f90:737), used in place of u, v.
template <typename T>
__global__ void __launch_bounds__(256)
psi_prep(const Statics<T>* sp, const T* u, const T* v, const double* und, const double* vnd, double* pu, double* pv) {
    const Statics<T>& a = *sp;
    const int Nx = a.Nx, Ny = a.Ny, P = a.P;
    const long n = (long)Nx * Ny;
    const long idx = (long)blockIdx.x * 256 + threadIdx.x;
    if (idx >= n) return;
    const int j = (int)(idx / Nx), i = (int)(idx - (long)j * Nx);
    const size_t off = (size_t)j * P + i;
    const int m = a.mask[off];
    const double ou = (m & OC_U) ? 1.0 : 0.0, ov = (m & OC_V) ? 1.0 : 0.0;
    const double uu = und ? und[off] : (double)u[off], vv = vnd ? vnd[off] : (double)v[off];
    const double hu = (double)fld(a.Hu, off, j, (m & OC_U) != 0), hv = (double)fld(a.Hv, off, j, (m & OC_V) != 0);
    if (a.barotropic) {
        pv[off] = j == 0 ? hv * ov * vv : hv * vv;
        pu[off] = i == Nx - 1 ? hu * ou * uu : hu * uu;
    } else {
        pv[off] = j == 0 ? ov * vv : vv;
        pu[off] = i == Nx - 1 ? ou * uu : uu;
    }
}

// Column sums sum_{q < j} pu(i, q), one thread per column in order of j; in
// column Nx they give psi(Nx, j) = -sum * A * dTheta (psi(Nx, 1) = 0).
template <typename T>
__global__ void __launch_bounds__(256)
psi_col(const Statics<T>* sp, const double* pu, double* colsum, double* psi) {
    const Statics<T>& a = *sp;
    const int Nx = a.Nx, Ny = a.Ny, P = a.P;
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= Nx) return;
    const double A = (double)a.A, dT = (double)a.dT;
    double s = 0.0;
    for (int j = 0; j < Ny; ++j) {
        const size_t off = (size_t)j * P + i;
        colsum[off] = s;
        if (i == Nx - 1) psi[off] = j == 0 ? 0.0 : -s * A * dT;
        s = s + pu[off];
    }
}

// Row sums sum_{q = i}^{Nx-1} pv(q, j) for 256 cells of one row per block,
// each cell adding q = i, i+1, ... in order (rows staged through LDS).  Row 1
// (j0 = 0) gives psi(i, 1); the interior combines both paths:
// psi = (-S_row*A*cos_v*dLambda + psi(Nx, j) - S_col*A*dTheta + psi(i, 1)) / 2.
template <typename T>
__global__ void __launch_bounds__(256)
psi_row(const Statics<T>* sp, int j0, const double* pv, const double* colsum, double* psi) {
    constexpr int CH = 2048;
    __shared__ double buf[CH];
    const Statics<T>& a = *sp;
    const int Nx = a.Nx, P = a.P;
    const int j = j0 + blockIdx.y;
    const int i0 = blockIdx.x * 256;
    const int i = i0 + threadIdx.x;
    const double* row = pv + (size_t)j * P;
    double s = 0.0;
    for (int c0 = i0; c0 < Nx - 1; c0 += CH) {
        const int len = min(CH, Nx - 1 - c0);
        __syncthreads();
        for (int q = threadIdx.x; q < len; q += 256) buf[q] = row[c0 + q];
        __syncthreads();
        const int qs = max(i - c0, 0);
        for (int q = qs; q < len; ++q) s = s + buf[q];
    }
    if (i >= Nx - 1) return;
    const double A = (double)a.A, dL = (double)a.dL, dT = (double)a.dT;
    const double cv = (double)a.cv[j];
    const size_t off = (size_t)j * P + i;
    if (j == 0) {
        psi[off] = -s * A * cv * dL;
    } else {
        psi[off] = (-s * A * cv * dL + psi[(size_t)j * P + Nx - 1] - colsum[off] * A * dT + psi[i]) / 2.0;
    }
}

template <typename T>
__global__ void __launch_bounds__(256)
sums_kernel(const Statics<T>* sp, int fn, const T* eta, const T* u, const T* v, const T* ch, const double* ce,
            double* partial) {
    const Statics<T>& a = *sp;
    const int Nx = a.Nx, Ny = a.Ny, P = a.P;
    double s[3] = {0.0, 0.0, 0.0}, c[3] = {0.0, 0.0, 0.0};
    const long n = (long)Nx * Ny;
    for (long idx = (long)blockIdx.x * 256 + threadIdx.x; idx < n; idx += (long)gridDim.x * 256) {
        const int j = (int)(idx / Nx), i = (int)(idx - (long)j * Nx);
        const size_t off = (size_t)j * P + i;
        if (!(a.mask[off] & OC_E)) continue;
        const int ip = i + 1 == Nx ? 0 : i + 1, jp = j + 1 == Ny ? 0 : j + 1;
        const double w = ce[j];
        const double e = (double)eta[off];
        const double H = (double)fld(a.He, off, j, true);
        const double D = fn ? e + H : H;
        const double uc = (double)u[off], ue = (double)u[(size_t)j * P + ip];
        const double vc = (double)v[off], vn = (double)v[(size_t)jp * P + i];
        const double ke = 0.5 * D * (0.5 * (uc * uc + ue * ue) + 0.5 * (vc * vc + vn * vn));
        const double pe = 0.5 * (double)a.G * (e * e);
        nsum_add(s[0], c[0], e * w);
        nsum_add(s[1], c[1], (ke + pe) * w);
        if (ch) nsum_add(s[2], c[2], (double)ch[off] * w);
    }
    __shared__ double rs[3][256], rc[3][256];
    const int tid = threadIdx.x;
    for (int k = 0; k < 3; ++k) { rs[k][tid] = s[k]; rc[k][tid] = c[k]; }
    __syncthreads();
    for (int stride = 128; stride > 0; stride >>= 1) {
        if (tid < stride)
            for (int k = 0; k < 3; ++k) {
                double ss = rs[k][tid], cc = rc[k][tid] + rc[k][tid + stride];
                nsum_add(ss, cc, rs[k][tid + stride]);
                rs[k][tid] = ss;
                rc[k][tid] = cc;
            }
        __syncthreads();
    }
    if (tid == 0)
        for (int k = 0; k < 3; ++k) {
            partial[(size_t)blockIdx.x * 6 + 2 * k] = rs[k][0];
            partial[(size_t)blockIdx.x * 6 + 2 * k + 1] = rc[k][0];
        }
}

__global__ void __launch_bounds__(256)
sums_final(const double* partial, int nblk, double* out) {
    double s[3] = {0.0, 0.0, 0.0}, c[3] = {0.0, 0.0, 0.0};
    const int tid = threadIdx.x;
    for (int b = tid; b < nblk; b += 256)
        for (int k = 0; k < 3; ++k) {
            c[k] += partial[(size_t)b * 6 + 2 * k + 1];
            nsum_add(s[k], c[k], partial[(size_t)b * 6 + 2 * k]);
        }
    __shared__ double rs[3][256], rc[3][256];
    for (int k = 0; k < 3; ++k) { rs[k][tid] = s[k]; rc[k][tid] = c[k]; }
    __syncthreads();
    for (int stride = 128; stride > 0; stride >>= 1) {
        if (tid < stride)
            for (int k = 0; k < 3; ++k) {
                double ss = rs[k][tid], cc = rc[k][tid] + rc[k][tid + stride];
                nsum_add(ss, cc, rs[k][tid + stride]);
                rs[k][tid] = ss;
                rc[k][tid] = cc;
            }
        __syncthreads();
    }
    if (tid == 0)
        for (int k = 0; k < 3; ++k) out[k] = rs[k][0] + rc[k][0];
}

}  // namespace swm
