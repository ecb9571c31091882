// engine.hip -- the batched FP64 HEOM RK4 propagator on CDNA4 (gfx950).
//
// Replaces the hot loops of twoD_spectrum_solver / dynamics_solver
// (etoheom-cpu/lib/dynamics.cpp:202-432) and their RHS total_ADO_dynamics[_Ht]
// (:53-200).  Design:
//
//  * Every instance (one propagation) owns a contiguous slice [M][N][N] of four
//    complex128 state buffers: Y (the state), ACC (the RK4 accumulator) and two
//    stage-input buffers S1/S2.  A step is four fused RHS+RK4 stage kernels:
//      stage 0: k1 = f(Y)   S1 = Y + h/2 k1
//      stage 1: k2 = f(S1)  ACC = Y + (S1 - Y)/3 + h/3 k2   S2 = Y + h/2 k2
//      stage 2: k3 = f(S2)  S1 = Y + h k3
//      stage 3: k4 = f(S1)  Y = ACC - Y/3 + S1/3 + h/6 k4
//    which is the reference's RK4 (dynamics.cpp:385-417) with V frozen over the
//    step (the h/6 k1 and h/3 k3 terms are rebuilt from the next stage's input
//    instead of written to ACC, which holds ACC - Y/3 so that stage 3 reads no Y:
//    7 reads + 5 writes of each ADO per step).  In the forward graph stage 0
//    also reads out Tr(rho_0 R) before the step and stage 3 builds V of the next
//    step (double-buffered by step parity); elsewhere a prologue kernel does both.
//  * f(rho)_n[a][b] = -i (V rho_n - rho_n V)[a][b] + damp_n rho_n[a][b]
//                     + sum_links ([a==s] w_row + [b==s] w_col) rho_m[a][b]
//    i.e. the one-hot S_j products of the reference reduce to element gathers of
//    the same (a, b) entry of the neighbour ADO.
//  * Small N (<= 16): one workgroup per tile of G = GE/N^2 ADOs, the tile and V
//    staged in LDS, GE/256 matrix elements per thread.  Larger N: one ADO per
//    workgroup with the commutator on FP64 MFMA tiles (N <= 32: 256 threads;
//    N <= 64: 512 threads), or from LDS on the VALU (stage_generic).
//  * The time loop is captured as a HIP graph of `chunk` steps; kernels read the
//    step index from a device counter that a one-thread kernel advances, so one
//    instantiated graph is replayed for the whole propagation.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <type_traits>
#include <vector>

#include "engine_impl.hpp"

namespace eh {

// ---------------------------------------------------------------- device data

// N <= 16: a tile of G = GE/N^2 ADOs per workgroup (GE state elements: 256, 512
// or 1024, the autotuner's choice), GE/256 elements per thread.  LATE: the update's state
// operand and the damping are loaded after the tile's barrier instead of with the input
// (GE = 512: 93 instead of 104 VGPRs, five waves/SIMD instead of four; measured 2.4 %
// faster on the 512-propagation grid, 3 % slower on one 462-ADO instance, whose launch
// the added load latency lengthens).
template <int NT, int GE, bool LATE = false>
__global__ __launch_bounds__(256) void stage_small(StageArgs A) {
    constexpr int NN = NT * NT;
    constexpr int G = NN >= GE ? 1 : GE / NN;
    constexpr int EPT = (G * NN + 255) / 256;
    __shared__ double2 sR[G * NN];
    __shared__ double2 sV[NN];
    extern __shared__ double2 sC[];   // G * 2 K N coupling rows/columns (launch-time size)
    const int tile = blockIdx.x;
    // a register copy: read once before any store (a reference would be re-read after them)
    const InstDev I = A.inst[A.tile_inst[tile]];
    const int step = *A.step_ctr + A.step_off;
    if (step >= I.nsteps) return;
    at_step(A, step);
    const int a0 = A.tile_ado0[tile], na = A.tile_nado[tile];
    const long long base = I.state_off + (long long)a0 * NN;
    const int nel = na * NN;
    // Every load below is unconditional (clamped address, selected result): the
    // neighbour indices and weights first, then own elements, y / acc, V, then the
    // neighbour gathers, which therefore wait for the indices alone.
    const int total = 2 * I.K * NT * na;
    const bool has_cpl = total > 0 && I.Km > 0;
    NbIdx<2, 2> nbx;
    if (has_cpl) gather_idx<2, 2, NT>(A, I, a0, na, threadIdx.x, 256, nbx);
    const bool vreal = A.vreal[I.step_off + step] != 0;
    // the one state operand the stage's update reads besides its input (Y in stages 1-2,
    // ACC in stage 3; stage 0 re-reads the just loaded, cached input)
    const int st = A.stage;
    const double2* wp = st == 3 ? A.acc : ((st == 1 || st == 2) ? A.y : A.rin);
    double2 rv[EPT], wv[EPT], gv[EPT];
#pragma unroll
    for (int q = 0; q < EPT; ++q) {
        const int e = threadIdx.x + q * 256;
        const int ec = e < nel ? e : nel - 1;
        const size_t gi = (size_t)(base + ec);
        rv[q] = A.rin[gi];
        if (!LATE) {
            wv[q] = wp[gi];
            gv[q] = A.damp[I.ado_off + a0 + ec / NN];
        }
    }
    // V into registers (written to LDS with the tile: a store right after its load
    // would wait for every load issued before it)
    constexpr int VPT = (NN + 255) / 256;
    double2 vv[VPT];
#pragma unroll
    for (int q = 0; q < VPT; ++q) {
        const int e = threadIdx.x + q * 256;
        vv[q] = A.V[I.v_off + (e < NN ? e : 0)];
    }
    const int cpn = 2 * I.K * NT;
    if (has_cpl) {
        gather_data<2, 2, NT>(A, I, na, threadIdx.x, 256, nbx, sC);
        if (total > 2 * 256 || I.Km > 2) gather_rest<2, 2>(A, I, a0, na, threadIdx.x, 256, sC);
    }
#pragma unroll
    for (int q = 0; q < VPT; ++q) {
        const int e = threadIdx.x + q * 256;
        if (e < NN) sV[e] = vv[q];
    }
#pragma unroll
    for (int q = 0; q < EPT; ++q) {
        const int e = threadIdx.x + q * 256;
        if (e < nel) sR[e] = rv[q];
    }
    __syncthreads();
    if (LATE) {
#pragma unroll
        for (int q = 0; q < EPT; ++q) {
            const int e = threadIdx.x + q * 256;
            const int ec = e < nel ? e : nel - 1;
            wv[q] = wp[(size_t)(base + ec)];
            gv[q] = A.damp[I.ado_off + a0 + ec / NN];
        }
    }
#pragma unroll
    for (int q = 0; q < EPT; ++q) {
        const int e = threadIdx.x + q * 256;
        if (e >= nel) break;
        const int nl = e / NN, ab = e - nl * NN, ra = ab / NT, cb = ab - ra * NT;
        const double2* R = sR + nl * NN;
        double cr = 0.0, ci = 0.0;   // (V rho - rho V)[ra][cb]
        if (vreal) {
#pragma unroll
            for (int c = 0; c < NT; ++c) {
                const double vac = sV[ra * NT + c].x, vcb = sV[c * NT + cb].x;
                const double2 rcb = R[c * NT + cb], rac = R[ra * NT + c];
                cr = fma(vac, rcb.x, fma(-rac.x, vcb, cr));
                ci = fma(vac, rcb.y, fma(-rac.y, vcb, ci));
            }
        } else {
#pragma unroll
            for (int c = 0; c < NT; ++c) {
                const double2 vac = sV[ra * NT + c], vcb = sV[c * NT + cb];
                const double2 rcb = R[c * NT + cb], rac = R[ra * NT + c];
                cr = fma(vac.x, rcb.x, fma(-vac.y, rcb.y, cr));
                ci = fma(vac.x, rcb.y, fma(vac.y, rcb.x, ci));
                cr = fma(-rac.x, vcb.x, fma(rac.y, vcb.y, cr));
                ci = fma(-rac.x, vcb.y, fma(-rac.y, vcb.x, ci));
            }
        }
        double dr = ci, di = -cr;   // -i * comm
        const double2 g = gv[q], r = rv[q];
        dr = fma(g.x, r.x, fma(-g.y, r.y, dr));
        di = fma(g.x, r.y, fma(g.y, r.x, di));
        const double2 cpl = coupling_of<NT>(I, sC + nl * cpn, ra, cb);
        dr += cpl.x;
        di += cpl.y;
        rk_store<true>(A, I, (size_t)(base + e), r, wv[q], wv[q], dr, di,
                 a0 + nl, ra, cb);
    }
    if (A.fused && a0 == 0) fused_tail(A, I, step, threadIdx.x, blockDim.x);
}

// any N <= 64: one ADO per workgroup, rho_n staged in dynamic LDS, V read through the cache
__global__ __launch_bounds__(256) void stage_generic(StageArgs A) {
    extern __shared__ double2 smem[];
    const int tile = blockIdx.x;
    // a register copy: read once before any store (a reference would be re-read after them)
    const InstDev I = A.inst[A.tile_inst[tile]];
    const int step = *A.step_ctr + A.step_off;
    if (step >= I.nsteps) return;
    at_step(A, step);
    const int N = I.N, NN = N * N;
    double2* sR = smem;
    double2* sC = smem + NN;
    const int a0 = A.tile_ado0[tile];
    const double2* V = A.V + I.v_off;
    const long long base = I.state_off + (long long)a0 * NN;
    for (int e = threadIdx.x; e < NN; e += blockDim.x) sR[e] = A.rin[base + e];
    gather_couplings(A, I, a0, 1, sC, threadIdx.x, blockDim.x);
    __syncthreads();
    const bool vreal = A.vreal[I.step_off + step] != 0;
    const double2 g = A.damp[I.ado_off + a0];
    for (int ab = threadIdx.x; ab < NN; ab += blockDim.x) {
        const int ra = ab / N, cb = ab - ra * N;
        double cr = 0.0, ci = 0.0;
        if (vreal) {
            for (int c = 0; c < N; ++c) {
                const double vac = V[ra * N + c].x, vcb = V[c * N + cb].x;
                const double2 rcb = sR[c * N + cb], rac = sR[ra * N + c];
                cr = fma(vac, rcb.x, fma(-rac.x, vcb, cr));
                ci = fma(vac, rcb.y, fma(-rac.y, vcb, ci));
            }
        } else {
            for (int c = 0; c < N; ++c) {
                const double2 vac = V[ra * N + c], vcb = V[c * N + cb];
                const double2 rcb = sR[c * N + cb], rac = sR[ra * N + c];
                cr = fma(vac.x, rcb.x, fma(-vac.y, rcb.y, cr));
                ci = fma(vac.x, rcb.y, fma(vac.y, rcb.x, ci));
                cr = fma(-rac.x, vcb.x, fma(rac.y, vcb.y, cr));
                ci = fma(-rac.x, vcb.y, fma(-rac.y, vcb.x, ci));
            }
        }
        double dr = ci, di = -cr;
        const double2 r = sR[ab];
        dr = fma(g.x, r.x, fma(-g.y, r.y, dr));
        di = fma(g.x, r.y, fma(g.y, r.x, di));
        const double2 cpl = coupling_of(I, sC, ra, cb);
        const double2 wv = rk_operand(A, (size_t)(base + ab), r);
        rk_store(A, I, (size_t)(base + ab), r, wv, wv, dr + cpl.x, di + cpl.y, a0, ra, cb);
    }
    if (A.fused && a0 == 0) fused_tail(A, I, step, threadIdx.x, blockDim.x);
}

// N in (16, NP]: one ADO per workgroup, the commutator on FP64 MFMA 16x16x4 tiles.
// rho_n (and V when VLDS) are staged in LDS as complex, zero-padded to NP x NP
// with row stride NP+1 (conflict-free for both operand patterns).  Each of the
// four waves owns (NP/16)^2/4 output tiles and accumulates
//   C = V rho - rho V  as  Re: Vr.rr - Vi.ri - rr.Vr + ri.Vi,  Im: Vr.ri + Vi.rr - rr.Vi - ri.Vr
// (the Vi terms only in steps with an active pulse).  MFMA operand layout:
// A[i][k] at lane i + 16k, B[k][j] at lane j + 16k, D[lane/16 + 4r][lane%16] in reg r.
// All global loads of the workgroup (rho_n, V, coupling gathers, and y / acc of
// the lane's output elements) are issued before the first wait.
// TP: two-phase gathers (neighbour indices and weights first, every load
// unconditional) -- a launch candidate offered when every instance of the group has
// 2 K N <= 512 and K_m <= 2.
template <int NP, bool VLDS, bool TP = false>
__global__ __launch_bounds__(256, NP == 32 ? 4 : 2) void stage_mfma(StageArgs A) {
    constexpr int LD = NP + 1;
    constexpr int NB = NP / 16;
    constexpr int TPW = NB * NB / 4;
    constexpr int SPT = NP * NP / 256;   // staged entries per thread
    extern __shared__ double2 smem[];
    double2* sR = smem;
    // V as real and imaginary planes of doubles: the 8-byte operand reads of the
    // commutator then walk a 33-double row stride without bank conflicts
    double* sVr = reinterpret_cast<double*>(smem + NP * LD);
    double* sVi = sVr + NP * LD;
    double2* sC = smem + (VLDS ? 2 : 1) * NP * LD;
    const int tile = blockIdx.x;
    // N <= 32: a register copy of the instance descriptor, read once before any store
    // (a reference would be re-read after them); N <= 64 keeps the reference: the
    // copy costs it half its occupancy
    using InstT = typename std::conditional<NP == 32, const InstDev, const InstDev&>::type;
    InstT I = A.inst[A.tile_inst[tile]];
    const int step = *A.step_ctr + A.step_off;
    if (step >= I.nsteps) return;
    at_step(A, step);
    const int N = I.N, NN = N * N;
    const int a0 = A.tile_ado0[tile];
    const long long base = I.state_off + (long long)a0 * NN;
    const double2* V = A.V + I.v_off;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int li = lane & 15, lk = lane >> 4;

    // Wave priority (two-phase variant): high while a wave issues its loads and stores, low
    // in the MFMA loop, so that with four workgroups per CU in different phases the
    // memory requests go out first (measured on the heptamer: 3.7 % faster over the first 300
    // steps, which hold the pulse windows and their complex V; 1.2 % over a 2000-step run).
    if (TP) __builtin_amdgcn_s_setprio(2);
    NbIdx<2, 2> nbx;
    if (TP) gather_idx<2, 2>(A, I, a0, 1, threadIdx.x, 256, nbx);
    double2 xr[SPT], xv[SPT];
    // y / acc of this lane's output elements.  The two-phase variant loads the update's
    // one state operand (Y in stages 1-2, ACC in stage 3) after the commutator instead,
    // into both arrays: 115 instead of 164 VGPRs, four workgroups per CU instead of three
    // (measured 6 % faster per RK4 step on the heptamer)
    double2 yv[TPW][4], av[TPW][4];
    if (TP) {   // every load unconditional: clamped address, selected result
#pragma unroll
        for (int q = 0; q < SPT; ++q) {
            const int e = threadIdx.x + q * 256;
            const int r = e / NP, c = e - r * NP;
            const bool in = r < N && c < N;
            const int ix = in ? r * N + c : 0;
            const double2 t = A.rin[base + ix];
            xr[q] = in ? t : make_double2(0.0, 0.0);
            if (VLDS) {
                const double2 u = V[ix];
                xv[q] = in ? u : make_double2(0.0, 0.0);
            }
        }
        gather_data<2, 2>(A, I, 1, threadIdx.x, 256, nbx, sC);
    } else {
#pragma unroll
        for (int q = 0; q < SPT; ++q) {
            const int e = threadIdx.x + q * 256;
            const int r = e / NP, c = e - r * NP;
            const bool in = r < N && c < N;
            xr[q] = in ? A.rin[base + r * N + c] : make_double2(0.0, 0.0);
            if (VLDS) xv[q] = in ? V[r * N + c] : make_double2(0.0, 0.0);
        }
#pragma unroll
        for (int t = 0; t < TPW; ++t) {
            const int tt = wave * TPW + t, ti = tt / NB, tj = tt - (tt / NB) * NB;
            const int col = tj * 16 + li;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = ti * 16 + lk + 4 * r;
                yv[t][r] = make_double2(0.0, 0.0);
                av[t][r] = make_double2(0.0, 0.0);
                if (row < N && col < N && A.stage != 0) {
                    const size_t gi = (size_t)(base + row * N + col);
                    if (A.stage < 3) yv[t][r] = A.y[gi];
                    if (A.stage == 3) av[t][r] = A.acc[gi];
                }
            }
        }
        gather_couplings<true>(A, I, a0, 1, sC, threadIdx.x, 256);
    }
#pragma unroll
    for (int q = 0; q < SPT; ++q) {
        const int e = threadIdx.x + q * 256;
        const int r = e / NP, c = e - r * NP;
        sR[r * LD + c] = xr[q];
        if (VLDS) {
            sVr[r * LD + c] = xv[q].x;
            sVi[r * LD + c] = xv[q].y;
        }
    }
    __syncthreads();
    if (TP) __builtin_amdgcn_s_setprio(0);   // the MFMA commutator
    const bool vreal = A.vreal[I.step_off + step] != 0;
    const double2 g = A.damp[I.ado_off + a0];
    const int kmax = (N + 3) & ~3;   // contraction over k < N, rounded up to the MFMA's 4
    auto vload = [&](int r, int c) -> double2 {
        if (VLDS) return make_double2(sVr[r * LD + c], sVi[r * LD + c]);
        return (r < N && c < N) ? V[r * N + c] : make_double2(0.0, 0.0);
    };
#pragma unroll
    for (int t = 0; t < TPW; ++t) {
        const int tt = wave * TPW + t;
        const int ti = tt / NB, tj = tt - (tt / NB) * NB;
        const int ar = ti * 16 + li;     // A-operand row
        const int bc = tj * 16 + li;     // B-operand column
        dbl4 cr = {0.0, 0.0, 0.0, 0.0}, ci = {0.0, 0.0, 0.0, 0.0};
        if (vreal) {
#pragma unroll 4
            for (int k0 = 0; k0 < NP; k0 += 4) {
                if (NP == 32 && k0 >= kmax) break;   // the padding beyond N holds zeros
                const int kk = k0 + lk;
                const double va = vload(ar, kk).x, vb = vload(kk, bc).x;
                const double2 rb = sR[kk * LD + bc], ra = sR[ar * LD + kk];
                cr = mfma4(va, rb.x, cr);
                ci = mfma4(va, rb.y, ci);
                cr = mfma4(ra.x, -vb, cr);
                ci = mfma4(ra.y, -vb, ci);
            }
        } else {
#pragma unroll 2
            for (int k0 = 0; k0 < NP; k0 += 4) {
                if (NP == 32 && k0 >= kmax) break;
                const int kk = k0 + lk;
                const double2 va = vload(ar, kk), vb = vload(kk, bc);
                const double2 rb = sR[kk * LD + bc], ra = sR[ar * LD + kk];
                cr = mfma4(va.x, rb.x, cr);
                ci = mfma4(va.x, rb.y, ci);
                cr = mfma4(-va.y, rb.y, cr);
                ci = mfma4(va.y, rb.x, ci);
                cr = mfma4(ra.x, -vb.x, cr);
                ci = mfma4(ra.y, -vb.x, ci);
                cr = mfma4(ra.y, vb.y, cr);
                ci = mfma4(ra.x, -vb.y, ci);
            }
        }
        const int col = tj * 16 + li;
        if (TP) __builtin_amdgcn_s_setprio(2);   // epilogue: operand loads, state stores
        if (TP) {   // the update's one state operand (clamped address), after the commutator
            const double2* wp = A.stage == 3 ? A.acc : ((A.stage == 1 || A.stage == 2) ? A.y : A.rin);
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = ti * 16 + lk + 4 * r;
                const double2 w = wp[(size_t)(base + (row < N && col < N ? row * N + col : 0))];
                yv[t][r] = w;
                av[t][r] = w;
            }
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int row = ti * 16 + lk + 4 * r;
            if (row < N && col < N) {
                double dr = ci[r], di = -cr[r];   // -i * C
                const double2 x = sR[row * LD + col];
                dr = fma(g.x, x.x, fma(-g.y, x.y, dr));
                di = fma(g.x, x.y, fma(g.y, x.x, di));
                const double2 cpl = coupling_of(I, sC, row, col);
                rk_store(A, I, (size_t)(base + row * N + col), x, yv[t][r], av[t][r], dr + cpl.x, di + cpl.y, a0,
                         row, col);
            }
        }
    }
    if (A.fused && a0 == 0) fused_tail(A, I, step, threadIdx.x, blockDim.x);
}


// N in (32, 64]: one ADO per workgroup of 512 threads (8 waves, each two 16x16
// output tiles of one tile row), two workgroups per CU so that one computes while
// the other loads.  rho_n sits in LDS (complex, stride 65); V comes from the
// instance's zero-padded 64x64 real / imaginary planes in global memory (shared by
// all the instance's workgroups, cache resident; the loads are unconditional); y and
// acc are loaded after the commutator to keep the registers of two workgroups.
// SYMV: for a real symmetric V (I.hsym, V = H outside the pulses) the A operand V[ar][kk]
// is read as V[kk][ar]: four rows of 16 consecutive doubles per wave instead of 16 rows
// of four (measured 1.2 % faster on the 256-propagation chain grid, same bits).
template <bool SYMV>
__device__ __forceinline__ void mfma64_tile(const StageArgs& A, int tile, int step) {
    constexpr int NP = 64, SPT = NP * NP / 512;
    extern __shared__ double2 smem[];
    const InstDev& I = A.inst[A.tile_inst[tile]];
    if (step >= I.nsteps) return;
    const int N = I.N, NN = N * N;
    const int LD = N + 1;   // rho_n in LDS at its own size (N x (N+1)): two workgroups fit a CU
    double2* sR = smem;
    double2* sC = smem + N * LD;
    const int a0 = A.tile_ado0[tile];
    const long long base = I.state_off + (long long)a0 * NN;
    const int st = A.stage;
    const bool vreal = A.vreal[I.step_off + step] != 0;
    const bool vsym = SYMV && vreal && I.hsym;
    const double2 g = A.damp[I.ado_off + a0];
    const double* Vr = A.Vpad + I.vp_off;
    const double* Vi = Vr + NP * NP;
    double2 xr[SPT];
#pragma unroll
    for (int q = 0; q < SPT; ++q) {
        const int e = threadIdx.x + q * 512;
        const int r = e >> 6, c = e & 63;
        const bool in = r < N && c < N;
        const double2 t = A.rin[base + (in ? r * N + c : 0)];
        xr[q] = in ? t : make_double2(0.0, 0.0);
    }
    gather_couplings<true>(A, I, a0, 1, sC, threadIdx.x, 512);
#pragma unroll
    for (int q = 0; q < SPT; ++q) {
        const int e = threadIdx.x + q * 512;
        const int r = e >> 6, c = e & 63;
        if (r < N && c < N) sR[r * LD + c] = xr[q];
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int li = lane & 15, lk = lane >> 4;
    const int ti = wave >> 1, tj0 = (wave & 1) * 2;
    const int ar = ti * 16 + li;
    dbl4 cr[2], ci[2];
#pragma unroll
    for (int t = 0; t < 2; ++t) {
        cr[t] = dbl4{0.0, 0.0, 0.0, 0.0};
        ci[t] = dbl4{0.0, 0.0, 0.0, 0.0};
    }
    // contraction over k < N only (rounded up to the MFMA's 4): the padding holds zeros
    const int kmax = (N + 3) & ~3;
    if (vreal) {
#pragma unroll 2
        for (int k0 = 0; k0 < kmax; k0 += 4) {
            const int kk = k0 + lk;
            const double va = vsym ? Vr[kk * NP + ar] : Vr[ar * NP + kk];
            const double2 ra = lds_at(sR, ar, kk, N, LD);
#pragma unroll
            for (int t = 0; t < 2; ++t) {
                const int bc = (tj0 + t) * 16 + li;
                const double vb = Vr[kk * NP + bc];
                const double2 rb = lds_at(sR, kk, bc, N, LD);
                cr[t] = mfma4(va, rb.x, cr[t]);
                ci[t] = mfma4(va, rb.y, ci[t]);
                cr[t] = mfma4(ra.x, -vb, cr[t]);
                ci[t] = mfma4(ra.y, -vb, ci[t]);
            }
        }
    } else {
#pragma unroll 2
        for (int k0 = 0; k0 < kmax; k0 += 4) {
            const int kk = k0 + lk;
            const double var = Vr[ar * NP + kk], vai = Vi[ar * NP + kk];
            const double2 ra = lds_at(sR, ar, kk, N, LD);
#pragma unroll
            for (int t = 0; t < 2; ++t) {
                const int bc = (tj0 + t) * 16 + li;
                const double vbr = Vr[kk * NP + bc], vbi = Vi[kk * NP + bc];
                const double2 rb = lds_at(sR, kk, bc, N, LD);
                cr[t] = mfma4(var, rb.x, cr[t]);
                ci[t] = mfma4(var, rb.y, ci[t]);
                cr[t] = mfma4(-vai, rb.y, cr[t]);
                ci[t] = mfma4(vai, rb.x, ci[t]);
                cr[t] = mfma4(ra.x, -vbr, cr[t]);
                ci[t] = mfma4(ra.y, -vbr, ci[t]);
                cr[t] = mfma4(ra.y, vbi, cr[t]);
                ci[t] = mfma4(ra.x, -vbi, ci[t]);
            }
        }
    }
    // y and acc of the lane's 8 output elements (unconditional, clamped), then the update
    const double2* wp = st == 3 ? A.acc : ((st == 1 || st == 2) ? A.y : A.rin);
    double2 wv[2][4];   // Y in stages 1-2, ACC in stage 3
#pragma unroll
    for (int t = 0; t < 2; ++t) {
        const int col = (tj0 + t) * 16 + li;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int row = ti * 16 + lk + 4 * r;
            const bool in = row < N && col < N;
            const size_t gi = (size_t)(base + (in ? row * N + col : 0));
            wv[t][r] = wp[gi];
        }
    }
#pragma unroll
    for (int t = 0; t < 2; ++t) {
        const int col = (tj0 + t) * 16 + li;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int row = ti * 16 + lk + 4 * r;
            if (row < N && col < N) {
                double dr = ci[t][r], di = -cr[t][r];   // -i * C
                const double2 x = sR[row * LD + col];
                dr = fma(g.x, x.x, fma(-g.y, x.y, dr));
                di = fma(g.x, x.y, fma(g.y, x.x, di));
                const double2 cpl = coupling_of(I, sC, row, col);
                // two operands: y (x in stage 0) and a (x in stage 1, ACC + x/3 in stage 3)
                const double2 w = wv[t][r];
                const double2 y0 = st == 0 ? x : w;
                const double2 a1 = st == 1 ? x : (st == 3 ? make_double2(fma(x.x, 1.0 / 3.0, w.x), fma(x.y, 1.0 / 3.0, w.y)) : w);
                rk_store<false, true>(A, I, (size_t)(base + row * N + col), x, y0, a1, dr + cpl.x, di + cpl.y, a0, row,
                                      col);
            }
        }
    }
    if (A.fused && a0 == 0) fused_tail(A, I, step, threadIdx.x, blockDim.x);
}

__global__ __launch_bounds__(512, 2) void stage_mfma64(StageArgs A) {
    const int step = *A.step_ctr + A.step_off;
    at_step(A, step);
    mfma64_tile<true>(A, blockIdx.x, step);
}

// ---------------------------------------------------------------------------------
// Sparse commutator stage for 32 < N <= 64.  V(t) = H - sum_q e_q(t) X'_q keeps the
// sparsity pattern of H and the three field operators over the whole run (made symmetric
// on the host: the pattern of column c is that of row c), and for aggregates it is thin:
// a double-exciton state couples to the doubles that share one site with it and to its
// two singles (the 10-site chain, N = 56: at most 13 entries per row, 7.4 on average).
// The commutator is then
//   (V rho)[a][b] - (rho V)[a][b] = sum_{c in P(a)} V[a][c] rho[c][b] - sum_{c in P(b)} rho[a][c] V[c][b],
// about 7x fewer flops than the dense product at N = 56, on the VALU (FP64 FMA runs at the
// rate of FP64 MFMA on gfx950, and MFMA work holds up the other waves' VALU work), so the
// launch is bound by its state traffic instead of by memory time plus MFMA time.
// One ADO per workgroup of 512 threads (two per CU): lane b of wave w computes column b
// of rows a = w, w + 8, ...  rho_n (stride N + 1) and the coupling rows/columns
// (gather_couplings) sit in LDS as in stage_mfma64, with each column's list: its row
// indices c and Re V[c][b].  V real and symmetric (no pulse on): row a's entries are
// column a's (LDS broadcasts).  Otherwise (the pulse windows) row a's values V[a][c] are
// scalar loads (a and c are uniform in the wave) and Im V[c][b] of the lane's column is
// loaded into registers once.  Summation order: row pattern ascending, then column
// pattern ascending (the exact zeros of the dense product are skipped; the sum differs
// from it by rounding only).
__global__ __launch_bounds__(512, 4) void stage_sparse(StageArgs A) {
    constexpr int CAP = kSpCap, PS = kSpCap + 1;
    extern __shared__ double2 smem[];
    const int step = *A.step_ctr + A.step_off;
    at_step(A, step);
    const int tile = blockIdx.x;
    const InstDev I = ro(A.inst, ro(A.tile_inst, tile));   // scalar loads, then registers
    if (step >= I.nsteps) return;
    const int N = I.N, NN = N * N, LD = N + 1, st = A.stage;
    const int a0 = ro(A.tile_ado0, tile);
    const long long base = I.state_off + (long long)a0 * NN;
    double2* sR = smem;                                              // rho_n, N x (N+1)
    double2* sC = sR + N * LD;                                       // coupling rows/columns, 2 K N
    double* sVr = reinterpret_cast<double*>(sC + 2 * I.K * N);       // [N][PS]: Re V[c][b]
    short* sIx = reinterpret_cast<short*>(sVr + N * PS);             // [N][PS]: c (-1 beyond)
    int* sCnt = reinterpret_cast<int*>(reinterpret_cast<char*>(sIx) + ((N * PS * 2 + 3) & ~3));   // [N]
    const int tid = threadIdx.x, b = tid & 63, w = tid >> 6;
    const double2* V = A.V + I.v_off;
    const long long pbase = (long long)I.sp_off * CAP;
    const bool vreal = A.vreal[I.step_off + step] != 0;
    const bool sym = vreal && I.hsym;
    // 1. the stage input, the column lists (N CAP <= 1024 entries: two per thread), the
    // coupling gathers; every load unconditional (clamped address, selected value)
    double2 xr[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        const int e = tid + q * 512, r = e >> 6, c = e & 63;
        const bool in = r < N && c < N;
        const double2 t = A.rin[base + (in ? r * N + c : 0)];
        xr[q] = in ? t : make_double2(0.0, 0.0);
    }
    int ix[2];
#pragma unroll
    for (int h = 0; h < 2; ++h) ix[h] = A.sp_idx[pbase + min(tid + h * 512, N * CAP - 1)];
    const int cnt = A.sp_cnt[I.sp_off + min(tid, N - 1)];
    gather_couplings<true>(A, I, a0, 1, sC, tid, 512);
    double vc[2];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int e = min(tid + h * 512, N * CAP - 1), c = e / CAP;
        vc[h] = V[(ix[h] >= 0 ? ix[h] : 0) * N + c].x;   // Re V[row][c] of column c's entry
    }
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        const int e = tid + q * 512, r = e >> 6, c = e & 63;
        if (r < N && c < N) sR[r * LD + c] = xr[q];
    }
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int e = tid + h * 512;
        if (e < N * CAP) {
            const int c = e / CAP, r = e - c * CAP;
            sIx[c * PS + r] = (short)ix[h];
            sVr[c * PS + r] = ix[h] >= 0 ? vc[h] : 0.0;
        }
    }
    if (tid < N) sCnt[tid] = cnt;
    __syncthreads();
    // 2. the Y / ACC operand of the lane's elements, then each row
    const double2 g = ro(A.damp, I.ado_off + a0);
    const double2* wp = st == 3 ? A.acc : A.y;
    const int bc = b < N ? b : 0;
    double2 wv[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        const int a = w + 8 * i;
        const bool in = a < N && b < N;
        wv[i] = st != 0 ? wp[base + (in ? a * N + b : 0)] : make_double2(0.0, 0.0);
    }
    // the wave's largest column count: the column loops skip the entries beyond it (a
    // uniform branch); column b's indices and values stay in registers (the same for
    // every row of the lane)
    int ncw = b < N ? sCnt[b] : 0;   // the pattern is symmetric: column b's count
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) ncw = max(ncw, __shfl_xor(ncw, o, 64));
    ncw = __builtin_amdgcn_readfirstlane(ncw);
    unsigned cix2[CAP / 2];   // two 16-bit indices per register
#pragma unroll
    for (int r = 0; r < CAP; r += 2) {
        const int c0 = sIx[bc * PS + r], c1 = sIx[bc * PS + r + 1];
        cix2[r / 2] = (unsigned)(c0 >= 0 ? c0 : 0) | ((unsigned)(c1 >= 0 ? c1 : 0) << 16);
    }
    auto cidx = [&](int r) { return (int)((cix2[r / 2] >> (16 * (r & 1))) & 0xffffu); };
    // Re V[c][b] (real symmetric V) or Im V[c][b] (otherwise) of column b's entries
    double cv[CAP];
#pragma unroll
    for (int r = 0; r < CAP; ++r)
        cv[r] = sym ? sVr[bc * PS + r] : (sIx[bc * PS + r] >= 0 ? V[cidx(r) * N + bc].y : 0.0);
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        const int a = w + 8 * i;
        if (a >= N) break;   // uniform in the wave
        const int nra = __builtin_amdgcn_readfirstlane(sCnt[a]);
        double2 acc = make_double2(0.0, 0.0);
        if (sym) {
            // (V rho)[a][b]: V real and symmetric, row a's entries are column a's (broadcasts)
#pragma unroll
            for (int r = 0; r < CAP; ++r) {
                if (r < nra) {
                    const int c = sIx[a * PS + r];
                    const double v = sVr[a * PS + r];
                    const double2 x = sR[c * LD + bc];
                    acc.x = fma(v, x.x, acc.x);
                    acc.y = fma(v, x.y, acc.y);
                }
            }
            // - (rho V)[a][b]: column b's entries
#pragma unroll
            for (int r = 0; r < CAP; ++r) {
                if (r < ncw) {
                    const double2 x = sR[a * LD + cidx(r)];
                    acc.x = fma(-x.x, cv[r], acc.x);
                    acc.y = fma(-x.y, cv[r], acc.y);
                }
            }
        } else {
            // (V rho)[a][b]: V[a][c] by scalar loads (row a and column c uniform)
#pragma unroll
            for (int r = 0; r < CAP; ++r) {
                if (r < nra) {
                    const int c = __builtin_amdgcn_readfirstlane((int)sIx[a * PS + r]);
                    acc = cmad(acc, ro(V, (long long)a * N + c), sR[c * LD + bc]);
                }
            }
            // - (rho V)[a][b]: column b's entries, V[c][b] = Re from LDS + i Im from registers
#pragma unroll
            for (int r = 0; r < CAP; ++r) {
                if (r < ncw) {
                    const double2 v = make_double2(-sVr[bc * PS + r], -cv[r]);
                    acc = cmad(acc, sR[a * LD + cidx(r)], v);
                }
            }
        }
        if (b < N) {
            const double2 x = sR[a * LD + b];
            double dr = acc.y, di = -acc.x;   // -i * acc
            dr = fma(g.x, x.x, fma(-g.y, x.y, dr));
            di = fma(g.x, x.y, fma(g.y, x.x, di));
            const double2 cpl = coupling_of(I, sC, a, b);
            const double2 ww = wv[i];
            const double2 y0 = st == 0 ? x : ww;
            const double2 a1 = st == 1 ? x : (st == 3 ? make_double2(fma(x.x, 1.0 / 3.0, ww.x), fma(x.y, 1.0 / 3.0, ww.y)) : ww);
            rk_store<true, true>(A, I, (size_t)(base + a * N + b), x, y0, a1, dr + cpl.x, di + cpl.y, a0, a, b);
        }
    }
    if (A.fused && a0 == 0) fused_tail(A, I, step, tid, 512);
}

// the value of v in lane l (l a compile-time constant) as a wave-uniform scalar
__device__ __forceinline__ double lane_d(double v, int l) {
    const long long u = __double_as_longlong(v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)u, l);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(u >> 32), l);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// LDS layout of one tile of stage_sparse_rows (N and K of its instance)
struct SxLds {
    double2* sR;    // rho_n, N x (N + 1)
    double2* sC;    // coupling rows/columns, 2 K N
    double* sVr;    // [N][kSpCap + 1]: Re V[a][c] of row a's entries (complex V: the waves' rows'
                    // V[a][c] and V[c][a], 2 kSpCap complex values per wave)
    short* sIx;     // [N][kSpCap + 1]: c (-1 beyond)
};
__device__ __forceinline__ SxLds sx_lds(double2* smem, int N, int K) {
    SxLds L;
    L.sR = smem;
    L.sC = L.sR + N * (N + 1);
    L.sVr = reinterpret_cast<double*>(L.sC + 2 * K * N);
    L.sIx = reinterpret_cast<short*>(L.sVr + N * (kSpCap + 1));
    return L;
}

// a tile's stage input and pattern lists into registers (every load unconditional: clamped
// address, selected value); sx_fill writes them to LDS
template <int R>
__device__ __forceinline__ void sx_load(const StageArgs& A, const InstDev& I, long long base, int tid, double2 (&xr)[R],
                                        int (&ix)[2], double (&vc)[2]) {
    constexpr int CAP = kSpCap;
    const int N = I.N;
    const double2* V = A.V + I.v_off;
    const long long pbase = (long long)I.sp_off * CAP;
#pragma unroll
    for (int q = 0; q < R; ++q) {   // rows 8q .. 8q + 7 (N <= 8R)
        const int e = tid + q * 512, r = e >> 6, c = e & 63;
        const bool in = r < N && c < N;
        const double2 t = A.rin[base + (in ? r * N + c : 0)];
        xr[q] = in ? t : make_double2(0.0, 0.0);
    }
#pragma unroll
    for (int h = 0; h < 2; ++h) ix[h] = A.sp_idx[pbase + min(tid + h * 512, N * CAP - 1)];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int e = min(tid + h * 512, N * CAP - 1), a = e / CAP;
        vc[h] = V[(long long)a * N + (ix[h] >= 0 ? ix[h] : 0)].x;   // Re V[a][c] of row a's entry
    }
}
template <int R>
__device__ __forceinline__ void sx_fill(const InstDev& I, const SxLds& L, int tid, const double2 (&xr)[R],
                                        const int (&ix)[2], const double (&vc)[2]) {
    constexpr int CAP = kSpCap, PS = kSpCap + 1;
    const int N = I.N, LD = N + 1;
#pragma unroll
    for (int q = 0; q < R; ++q) {
        const int e = tid + q * 512, r = e >> 6, c = e & 63;
        if (r < N && c < N) L.sR[r * LD + c] = xr[q];
    }
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int e = tid + h * 512;
        if (e < N * CAP) {
            const int a = e / CAP, r = e - a * CAP;
            L.sIx[a * PS + r] = (short)ix[h];
            L.sVr[a * PS + r] = ix[h] >= 0 ? vc[h] : 0.0;
        }
    }
}

// The rows of one tile: the partial outputs p1 (out) without the (rho V) term, and
// (rho V)[b][a] in tq.
template <int S, int R>
__device__ __forceinline__ void sx_rows(const StageArgs& A, const InstDev& I, int a0, long long base, const SxLds& L,
                                        bool sym, double2 (&wv)[R], double2 (&p1)[R], double2 (&tq)[R]) {
    constexpr int CAP = kSpCap, PS = kSpCap + 1;
    const int N = I.N, LD = N + 1;
    // the wave index (and so every row a = w + 8 i) as a scalar: row offsets go to the SALU
    // (stage 1 measured slower with it)
    const int tid = threadIdx.x, b = tid & 63, w = S == 1 ? tid >> 6 : __builtin_amdgcn_readfirstlane(tid >> 6);
    // idle lanes (b >= N) read what the first lane of their 16-lane group reads: a 128-bit LDS
    // read is served 16 lanes at a time, and element 0 of a row or column would share banks
    // with the group's own elements (a 2-way conflict in every entry's reads at N = 56)
    const int bc = b < N ? b : ((b & ~15) < N ? (b & ~15) : 0), l16 = b & (CAP - 1);
    const double2* V = A.V + I.v_off;
    // the state operand of the lane's elements (Y in stage 2, ACC in stage 3; stage 1 loads Y
    // after the rows)
    if (S >= 2) {
        const double2* wp = S == 3 ? A.acc : A.y;
#pragma unroll
        for (int i = 0; i < R; ++i) {
            const int a = w + 8 * i;
            wv[i] = wp[base + (a < N && b < N ? a * N + b : 0)];
        }
    }
    __syncthreads();   // the tile's input, lists and couplings are in LDS
    const double2 g = ro(A.damp, I.ado_off + a0);
    const double hc = S == 0 || S == 1 ? I.h2 : (S == 2 ? I.h : I.hb[3]);
#pragma unroll
    for (int i = 0; i < R; ++i) {
        const int a = w + 8 * i;
        if (a >= N) break;   // uniform in the wave
        // row a's entries, lane r holding entry r: its column c and Re V[a][c] (vs, from the
        // LDS list), or V[a][c] (u) and V[c][a] (d) from the step's V
        const int cr = L.sIx[a * PS + l16];
        const int nra = __popcll(__ballot(b < CAP && cr >= 0));
        double2 ar = make_double2(0.0, 0.0), at = ar;
        if (sym) {
            const double vs = L.sVr[a * PS + l16];
#pragma unroll 1
            for (int r = 0; r < nra; ++r) {
                const int c = __builtin_amdgcn_readlane(cr, r);
                const double2 xa = L.sR[c * LD + bc], xb = L.sR[bc * LD + c];
                const double v = lane_d(vs, r);
                ar.x = fma(v, xa.x, ar.x);
                ar.y = fma(v, xa.y, ar.y);
                at.x = fma(xb.x, v, at.x);
                at.y = fma(xb.y, v, at.y);
            }
        } else {
            // V[a][c] and V[c][a] of the row's entries into this wave's slots (the Re V list
            // is not read with complex V), then read back as LDS broadcasts
            const int cs = cr >= 0 ? cr : 0;   // a, cs < N: 24-bit products
            const double2 u = V[__umul24(a, N) + cs], d = V[__umul24(cs, N) + a];
            double2* slot = reinterpret_cast<double2*>(L.sVr) + w * 2 * CAP;
            if (b < CAP) {
                slot[b] = u;
                slot[CAP + b] = d;
            }
            // the wave reads what its own lanes wrote: a wave's LDS operations complete in
            // issue order, so no barrier; this keeps the compiler from moving them
            __builtin_amdgcn_wave_barrier();
#pragma unroll 1
            for (int r = 0; r < nra; ++r) {
                const int c = __builtin_amdgcn_readlane(cr, r);
                const double2 xa = L.sR[c * LD + bc], xb = L.sR[bc * LD + c];
                ar = cmad(ar, slot[r], xa);
                at = cmad(at, xb, slot[CAP + r]);
            }
        }
        tq[i] = at;
        // fold the damping, the coupling terms and the state operand in
        const double2 x = L.sR[a * LD + bc];
        double kr = ar.y, ki = -ar.x;   // -i * (V rho)
        kr = fma(g.x, x.x, fma(-g.y, x.y, kr));
        ki = fma(g.x, x.y, fma(g.y, x.x, ki));
        const double2 cpl = coupling_of(I, L.sC, a, bc);
        kr += cpl.x;
        ki += cpl.y;
        if (S == 1) {   // stage 1 keeps the partial derivative: its two outputs follow the exchange
            p1[i] = make_double2(kr, ki);
        } else {
            double2 o = S == 0 ? x : wv[i];   // out = o + hc k
            if (S == 3) o = make_double2(fma(x.x, 1.0 / 3.0, o.x), fma(x.y, 1.0 / 3.0, o.y));
            p1[i] = make_double2(fma(hc, kr, o.x), fma(hc, ki, o.y));
        }
    }
}

// Sparse commutator stage, row-pattern form (one instantiation per RK4 stage S and rows per
// lane R).  Thread (lane b, row a) runs over row a's pattern P(a) (wave-uniform) once and
// forms both
//   (V rho)[a][b] = sum_{c in P(a)} V[a][c] rho[c][b]   (row c of rho: consecutive in the wave)
//   (rho V)[b][a] = sum_{c in P(a)} rho[b][c] V[c][a]   (column c of rho: stride N + 1, odd,
//                                                        so the wave's reads hit distinct banks)
// (the pattern is symmetric, so P(a) is also column a's pattern).  Everything of the update
// but the (rho V)[a][b] term goes into the element's partial outputs at once (the damping,
// the coupling terms, the state operand); once every row is done the (rho V) values are
// swapped through rho_n's LDS (element [b][a] to the thread that owns it) and completed.
// Row a's entries sit one per lane in lanes 0..15: the column c reaches the wave by
// readlane.  V real and symmetric: Re V[a][c] from the LDS list, also by readlane.
// Otherwise V[a][c] and V[c][a] are loaded from the step's V, written to the wave's slots
// in the LDS of that list (not read then), and read back as LDS broadcasts.  Summation
// order: each sum over P(a) ascending (the exact zeros of the dense product are skipped; the
// sum differs from it by rounding only).
template <int S, int R>
__global__ __launch_bounds__(512, 4) void stage_sparse_rows(StageArgs A) {
    extern __shared__ double2 smem[];
    const int step = *A.step_ctr + A.step_off;
    at_step(A, step);
    const int tile = blockIdx.x, tid = threadIdx.x, b = tid & 63;
    const int w = S == 1 ? tid >> 6 : __builtin_amdgcn_readfirstlane(tid >> 6);   // as in sx_rows
    const InstDev I = ro(A.inst, ro(A.tile_inst, tile));   // scalar loads, then registers
    if (step >= I.nsteps) return;
    const int N = I.N, LD = N + 1;
    const int a0 = ro(A.tile_ado0, tile);
    const long long base = I.state_off + (long long)a0 * N * N;
    const SxLds L = sx_lds(smem, N, I.K);
    if (S != 1) {
        // the input, the lists and the coupling gathers' tables in flight together; the
        // input into LDS (its registers free), then the neighbour rows/columns (two-phase
        // gathers: three entries and two modes per thread, any more by gather_rest).
        // Measured on 256 propagations of the 10-site chain (N = 56, 1771 ADOs each): stage 0
        // 14.1 -> 13.3 ms, stages 2-3 unchanged; stage 1,
        // whose rows hold more registers, spills with it and keeps the one-phase gathers.
        double2 xr[R];
        int ix[2];
        double vc[2];
        NbIdx<3, 2> nbx;
        sx_load<R>(A, I, base, tid, xr, ix, vc);
        gather_idx<3, 2>(A, I, a0, 1, tid, 512, nbx);
        sx_fill<R>(I, L, tid, xr, ix, vc);
        gather_data<3, 2>(A, I, 1, tid, 512, nbx, L.sC);
        if (2 * I.K * N > 3 * 512 || I.Km > 2) gather_rest<3, 2>(A, I, a0, 1, tid, 512, L.sC);
    } else {
        double2 xr[R];
        int ix[2];
        double vc[2];
        sx_load<R>(A, I, base, tid, xr, ix, vc);
        gather_couplings(A, I, a0, 1, L.sC, tid, 512);
        sx_fill<R>(I, L, tid, xr, ix, vc);
    }
    double2 wv[R], p1[R], tq[R];
    sx_rows<S, R>(A, I, a0, base, L, A.vreal[I.step_off + step] != 0 && I.hsym, wv, p1, tq);
    __syncthreads();   // every read of rho_n, the couplings and the lists is done
    if (S == 1) {
        // two outputs, out = Y + h2 k and ACC = (x + Y) / 3 + hb[1] k, from k = k' + i (rho V):
        // x stays in rho_n's LDS, and the (rho V) values go through the LDS of the couplings
        // and lists in two rounds of H rows (the launch holds H (N | 1) elements there)
        const int H = (N + 1) / 2, LE = N | 1;
        double2* sE = L.sC;
#pragma unroll
        for (int i = 0; i < R; ++i) {
            const int a = w + 8 * i;
            wv[i] = A.y[base + (a < N && b < N ? a * N + b : 0)];
        }
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            if (h) __syncthreads();   // the first round's values are read
            if (b >= h * H && b < N && b < (h + 1) * H) {
#pragma unroll
                for (int i = 0; i < R; ++i) {
                    const int a = w + 8 * i;
                    if (a < N) sE[(b - h * H) * LE + a] = tq[i];
                }
            }
            __syncthreads();
#pragma unroll
            for (int i = 0; i < R; ++i) {
                const int a = w + 8 * i;
                if (a >= h * H && a < N && a < (h + 1) * H && b < N) {   // a: uniform in the wave
                    const double2 t = sE[(a - h * H) * LE + b];   // (rho V)[a][b]
                    const double2 x = L.sR[a * LD + b], y = wv[i];
                    const double kr = p1[i].x - t.y, ki = p1[i].y + t.x;   // k' + i (rho V)
                    const size_t gi = (size_t)(base + a * N + b);
                    const double2 q = make_double2((x.x + y.x) * (1.0 / 3.0), (x.y + y.y) * (1.0 / 3.0));
                    st_state<true>(A, A.acc + gi, make_double2(fma(I.hb[1], kr, q.x), fma(I.hb[1], ki, q.y)));
                    st_state<true>(A, A.out + gi, make_double2(fma(I.h2, kr, y.x), fma(I.h2, ki, y.y)));
                }
            }
        }
        if (A.fused && a0 == 0) fused_tail(A, I, step, tid, 512);
        return;
    }
    // (rho V)[b][a] into rho_n's place [b][a], then each thread's own (rho V)[a][b]:
    // out = p1 + hc i (rho V)
    if (b < N) {
#pragma unroll
        for (int i = 0; i < R; ++i) {
            const int a = w + 8 * i;
            if (a < N) L.sR[b * LD + a] = tq[i];
        }
    }
    __syncthreads();
    const double hc = S == 0 ? I.h2 : (S == 2 ? I.h : I.hb[3]);
#pragma unroll
    for (int i = 0; i < R; ++i) {
        const int a = w + 8 * i;
        if (a < N && b < N) {
            const double2 t = L.sR[a * LD + b];
            const size_t gi = (size_t)(base + a * N + b);
            st_state<true>(A, A.out + gi, make_double2(fma(-hc, t.y, p1[i].x), fma(hc, t.x, p1[i].y)));
        }
    }
    if (A.fused && a0 == 0) fused_tail(A, I, step, tid, 512);
}

// The kernel of stage `stage` in the row-pattern candidate.  Stage 1 stores two outputs
// (out and ACC) and exchanges its (rho V) values through the LDS past rho_n, in two rounds;
// where that LDS is too small for half of them (rows1 false) it runs the column-list kernel.
typedef void (*StageKernel)(StageArgs);
static StageKernel sparse_rows_kernel(int stage, int rows, bool rows1) {
    static const StageKernel k[2][4] = {
        {stage_sparse_rows<0, 7>, stage_sparse_rows<1, 7>, stage_sparse_rows<2, 7>, stage_sparse_rows<3, 7>},
        {stage_sparse_rows<0, 8>, stage_sparse_rows<1, 8>, stage_sparse_rows<2, 8>, stage_sparse_rows<3, 8>}};
    if ((stage & 3) == 1 && !rows1) return stage_sparse;
    return k[rows == 7 ? 0 : 1][stage & 3];
}
// rows per lane of the row-pattern kernel for a group (32 < N <= 64) whose largest N^2 is maxnn
static int sparse_rows_of(int maxnn) { return maxnn <= 56 * 56 ? 7 : 8; }

// Coupling term of output element (row, col) of the ADO whose table row starts at lrow,
// gathered straight from the stage input: the same element of the up / down neighbours
// of the bath whose site is row (weights i x, i alpha y) and of the bath whose site is
// col (-i x, -i alpha~ y), summed in gather_couplings' order (modes k, up then down;
// the row term, then the column term).
__device__ __forceinline__ double2 coupling_direct(const StageArgs& A, const InstDev& I, long long lrow, int row,
                                                   int col) {
    const int N = I.N, K = I.K, Km = I.Km;
    const long long e = (long long)row * N + col;
    const double2* rb = A.rin + I.state_off;
    double2 c = make_double2(0.0, 0.0);
#pragma unroll
    for (int rc = 0; rc < 2; ++rc) {
        const int j = (rc == 0 ? row : col) - I.so;
        if (j < 0 || j >= K) continue;
        const double sg = rc == 0 ? A.gsign : -A.gsign;
        double2 acc = make_double2(0.0, 0.0);
        for (int k = 0; k < Km; ++k) {
            const long long q = lrow + j + k * K;
            const int mu = A.nb_up[q], md = A.nb_dn[q];
            const double2 xu = mu >= 0 ? rb[(long long)mu * N * N + e] : make_double2(0.0, 0.0);
            const double2 xd = md >= 0 ? rb[(long long)md * N * N + e] : make_double2(0.0, 0.0);
            const double x = A.cx[q], y = A.cy[q];
            const double2 al = rc == 0 ? A.alpha[I.al_off + j * Km + k] : A.alpha_t[I.al_off + j * Km + k];
            acc = cmad(acc, make_double2(0.0, sg * x), xu);
            acc = cmad(acc, make_double2(-sg * al.y * y, sg * al.x * y), xd);
        }
        c.x += acc.x;
        c.y += acc.y;
    }
    return c;
}

// N > 64: one ADO per gridDim.y workgroups of 256 threads and nothing in LDS (rho_n alone
// is 256 KB at N = 128).  The 16 x 16 output tiles of the ADO are dealt round-robin to the
// 4 gridDim.y waves; the MFMA operands (V and rho_n, zero beyond N) are loaded straight
// from global memory through the caches, and each output element's coupling term is
// gathered from the neighbours' same element (coupling_direct).
__global__ __launch_bounds__(256) void stage_mfma_wide(StageArgs A) {
    const int tile = blockIdx.x;
    const InstDev I = A.inst[A.tile_inst[tile]];   // register copy, read before any store
    const int step = *A.step_ctr + A.step_off;
    if (step >= I.nsteps) return;
    at_step(A, step);
    const int N = I.N, NN = N * N;
    const int a0 = A.tile_ado0[tile];
    const long long base = I.state_off + (long long)a0 * NN;
    const double2* R = A.rin + base;
    const double2* V = A.V + I.v_off;
    const bool vreal = A.vreal[I.step_off + step] != 0;
    const double2 g = A.damp[I.ado_off + a0];
    const long long lrow = I.lk_off + (long long)a0 * I.d;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int li = lane & 15, lk = lane >> 4;
    const int T = (N + 15) >> 4, kmax = (N + 3) & ~3;
    const double2 z = make_double2(0.0, 0.0);
    for (int tt = (int)blockIdx.y * 4 + wave; tt < T * T; tt += 4 * (int)gridDim.y) {
        const int ti = tt / T, tj = tt - ti * T;
        const int ar = ti * 16 + li, bc = tj * 16 + li;   // A-operand row, B-operand column
        dbl4 cr = {0.0, 0.0, 0.0, 0.0}, ci = {0.0, 0.0, 0.0, 0.0};
        if (vreal) {
#pragma unroll 2
            for (int k0 = 0; k0 < kmax; k0 += 4) {
                const int kk = k0 + lk;
                const bool ain = ar < N && kk < N, bin = kk < N && bc < N;
                const int ia = ain ? ar * N + kk : 0, ib = bin ? kk * N + bc : 0;   // (ar, kk), (kk, bc)
                const double va0 = V[ia].x, vb0 = V[ib].x;
                const double2 ra0 = R[ia], rb0 = R[ib];
                const double va = ain ? va0 : 0.0, vb = bin ? vb0 : 0.0;
                const double2 ra = ain ? ra0 : z, rb = bin ? rb0 : z;
                cr = mfma4(va, rb.x, cr);
                ci = mfma4(va, rb.y, ci);
                cr = mfma4(ra.x, -vb, cr);
                ci = mfma4(ra.y, -vb, ci);
            }
        } else {
#pragma unroll 2
            for (int k0 = 0; k0 < kmax; k0 += 4) {
                const int kk = k0 + lk;
                const bool ain = ar < N && kk < N, bin = kk < N && bc < N;
                const int ia = ain ? ar * N + kk : 0, ib = bin ? kk * N + bc : 0;
                const double2 va0 = V[ia], vb0 = V[ib], ra0 = R[ia], rb0 = R[ib];
                const double2 va = ain ? va0 : z, vb = bin ? vb0 : z;
                const double2 ra = ain ? ra0 : z, rb = bin ? rb0 : z;
                cr = mfma4(va.x, rb.x, cr);
                ci = mfma4(va.x, rb.y, ci);
                cr = mfma4(-va.y, rb.y, cr);
                ci = mfma4(va.y, rb.x, ci);
                cr = mfma4(ra.x, -vb.x, cr);
                ci = mfma4(ra.y, -vb.x, ci);
                cr = mfma4(ra.y, vb.y, cr);
                ci = mfma4(ra.x, -vb.y, ci);
            }
        }
        const int col = tj * 16 + li;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int row = ti * 16 + lk + 4 * r;
            if (row < N && col < N) {
                const size_t gi = (size_t)(base + row * N + col);
                double dr = ci[r], di = -cr[r];   // -i * C
                const double2 x = A.rin[gi];
                dr = fma(g.x, x.x, fma(-g.y, x.y, dr));
                di = fma(g.x, x.y, fma(g.y, x.x, di));
                const double2 cpl = coupling_direct(A, I, lrow, row, col);
                const double2 w = rk_operand(A, gi, x);
                rk_store(A, I, gi, x, w, w, dr + cpl.x, di + cpl.y, a0, row, col);
            }
        }
    }
    if (A.fused && a0 == 0 && blockIdx.y == 0) fused_tail(A, I, step, threadIdx.x, blockDim.x);
}

// rho_0 = |0><0| for every instance, all other ADOs zero; step counter to 0
// Persistent, software-pipelined variant of stage_mfma for N <= 32 with at most
// KM modes per bath and 2 K N <= 512: each workgroup walks the tiles
// t = blockIdx.x, +gridDim.x, ... and, while its four waves run the MFMA
// commutator and the epilogue of tile t from LDS buffer b, the loads of tile
// t + gridDim.x are already in flight into registers (its neighbour indices were
// loaded one tile earlier still).  Vector loads retire in order, so every small
// load the loop body waits on (tile descriptors, instance fields, the V-is-real
// flag, the damping) is issued before the prefetch and kept in registers; the
// only waits inside the loop are then the commit of the prefetched tile and the
// descriptors of the tile two ahead.  rho and the coupling rows/columns are
// double-buffered in LDS; V is held as real and imaginary planes of doubles (row
// stride NP+1: conflict-free for the 8-byte operand reads) and restaged only when
// the tile's instance changes.
namespace pipe {
constexpr int NP = 32, LD = NP + 1, SPT = NP * NP / 256, ENT = 2;

struct Tile {   // a tile's descriptor and its instance's fields, in registers
    int inst, a0, N, K, Km, d, so, ado_off, al_off, v_off, step_off, nsteps;
    long long state_off, lk_off;
};

__device__ __forceinline__ Tile tile_of(const StageArgs& A, int t) {
    Tile T;
    T.inst = A.tile_inst[t];
    T.a0 = A.tile_ado0[t];
    const InstDev& I = A.inst[T.inst];
    T.N = I.N;
    T.K = I.K;
    T.Km = I.Km;
    T.d = I.d;
    T.so = I.so;
    T.ado_off = I.ado_off;
    T.al_off = I.al_off;
    T.v_off = I.v_off;
    T.step_off = I.step_off;
    T.nsteps = I.nsteps;
    T.state_off = I.state_off;
    T.lk_off = I.lk_off;
    return T;
}

template <int KM>
struct Idx {   // neighbour indices of a tile, loaded one tile ahead of its data
    int up[ENT][KM], dn[ENT][KM];
};
template <int KM>
struct Dat {   // everything else of a tile
    double2 x[SPT];
    double2 y[4], a[4];
    double2 gu[ENT][KM], gd[ENT][KM];
    double cx[ENT][KM], cy[ENT][KM];
    double2 al[ENT][KM];
};

// entry e of the coupling rows/columns -> (bath j, row/column, element)
__device__ __forceinline__ void entry_of(int e, int N, int so, int& j, int& rc, int& ab) {
    j = e / (2 * N);
    const int rem = e - j * 2 * N;
    rc = rem / N;
    const int pos = rem - rc * N;
    const int s = j + so;
    ab = rc == 0 ? s * N + pos : pos * N + s;
}

template <int KM>
__device__ __forceinline__ void load_idx(const StageArgs& A, const Tile& T, Idx<KM>& X) {
    const int per = 2 * T.K * T.N;
#pragma unroll
    for (int r = 0; r < ENT; ++r) {
        const int e = threadIdx.x + 256 * r;
        int j = 0, rc = 0, ab = 0;
        if (e < per) entry_of(e, T.N, T.so, j, rc, ab);
#pragma unroll
        for (int k = 0; k < KM; ++k) {
            X.up[r][k] = -1;
            X.dn[r][k] = -1;
            if (e < per && k < T.Km) {
                const long long q = T.lk_off + (long long)T.a0 * T.d + j + k * T.K;
                X.up[r][k] = A.nb_up[q];
                X.dn[r][k] = A.nb_dn[q];
            }
        }
    }
}

template <int KM>
__device__ __forceinline__ void load_dat(const StageArgs& A, const Tile& T, const Idx<KM>& X, Dat<KM>& D) {
    const int N = T.N, NN = N * N;
    const long long base = T.state_off + (long long)T.a0 * NN;
#pragma unroll
    for (int q = 0; q < SPT; ++q) {
        const int e = threadIdx.x + q * 256;
        const int r = e / NP, c = e - r * NP;
        D.x[q] = (r < N && c < N) ? A.rin[base + r * N + c] : make_double2(0.0, 0.0);
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int ti = wave >> 1, tj = wave & 1, li = lane & 15, lk = lane >> 4;
    const int col = tj * 16 + li;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int row = ti * 16 + lk + 4 * r;
        D.y[r] = make_double2(0.0, 0.0);
        D.a[r] = make_double2(0.0, 0.0);
        if (row < N && col < N && A.stage != 0) {
            const size_t gi = (size_t)(base + row * N + col);
            if (A.stage < 3) D.y[r] = A.y[gi];
            if (A.stage == 3) D.a[r] = A.acc[gi];
        }
    }
    const int per = 2 * T.K * N;
#pragma unroll
    for (int r = 0; r < ENT; ++r) {
        const int e = threadIdx.x + 256 * r;
        int j = 0, rc = 0, ab = 0;
        if (e < per) entry_of(e, N, T.so, j, rc, ab);
#pragma unroll
        for (int k = 0; k < KM; ++k) {
            const int mu = X.up[r][k], md = X.dn[r][k];
            D.gu[r][k] = mu >= 0 ? A.rin[T.state_off + (long long)mu * NN + ab] : make_double2(0.0, 0.0);
            D.gd[r][k] = md >= 0 ? A.rin[T.state_off + (long long)md * NN + ab] : make_double2(0.0, 0.0);
            D.cx[r][k] = 0.0;
            D.cy[r][k] = 0.0;
            D.al[r][k] = make_double2(0.0, 0.0);
            if (e < per && k < T.Km) {
                const long long q = T.lk_off + (long long)T.a0 * T.d + j + k * T.K;
                D.cx[r][k] = A.cx[q];
                D.cy[r][k] = A.cy[q];
                D.al[r][k] = rc == 0 ? A.alpha[T.al_off + j * T.Km + k] : A.alpha_t[T.al_off + j * T.Km + k];
            }
        }
    }
}
}  // namespace pipe

template <int KM>
__global__ __launch_bounds__(256, 2) void stage_mfma_pipe(StageArgs A, int ntiles, int cpn, double h2, double h,
                                                          double hb) {
    using namespace pipe;
    extern __shared__ double2 smem[];
    double2* const sR0 = smem;
    double2* const sR1 = smem + NP * LD;
    double* const sVr = reinterpret_cast<double*>(smem + 2 * NP * LD);
    double* const sVi = sVr + NP * LD;
    double2* const sC0 = reinterpret_cast<double2*>(sVi + NP * LD);
    double2* const sC1 = sC0 + cpn;
    const int step = *A.step_ctr + A.step_off;
    at_step(A, step);
    int t = blockIdx.x;
    if (t >= ntiles) return;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int li = lane & 15, lk = lane >> 4;
    const int ti = wave >> 1, tj = wave & 1;   // one 16x16 output tile per wave
    const double gs = A.gsign;
    const int st = A.stage;

    Idx<KM> ix;
    Dat<KM> dt;
    Tile cur = tile_of(A, t);
    load_idx(A, cur, ix);
    load_dat(A, cur, ix, dt);
    int tn = t + gridDim.x;
    Tile nxt = cur;
    if (tn < ntiles) {
        nxt = tile_of(A, tn);
        load_idx(A, nxt, ix);
    }
    int vinst = -1;
    int b = 0;
    while (true) {
        const int N = cur.N;
        const long long base = cur.state_off + (long long)cur.a0 * (N * N);
        const bool live = step < cur.nsteps;
        if (cur.inst != vinst) {   // (re)stage V for this instance
            __syncthreads();
            const double2* V = A.V + cur.v_off;
            for (int e = threadIdx.x; e < NP * NP; e += 256) {
                const int r = e / NP, c = e - r * NP;
                const double2 v = (r < N && c < N) ? V[r * N + c] : make_double2(0.0, 0.0);
                sVr[r * LD + c] = v.x;
                sVi[r * LD + c] = v.y;
            }
            vinst = cur.inst;
        }
        const bool vreal = A.vreal[cur.step_off + (live ? step : 0)] != 0;
        const double2 g = A.damp[cur.ado_off + cur.a0];
        // commit tile `cur` into LDS buffer b
        double2* const Rb = b ? sR1 : sR0;
        double2* const Cb = b ? sC1 : sC0;
#pragma unroll
        for (int q = 0; q < SPT; ++q) {
            const int e = threadIdx.x + q * 256;
            const int r = e / NP, c = e - r * NP;
            Rb[r * LD + c] = dt.x[q];
        }
        {
            const int per = 2 * cur.K * N;
#pragma unroll
            for (int r = 0; r < ENT; ++r) {
                const int e = threadIdx.x + 256 * r;
                if (e < per) {
                    const int rc = (e - (e / (2 * N)) * 2 * N) / N;
                    const double sg = rc == 0 ? gs : -gs;
                    double2 acc = make_double2(0.0, 0.0);
#pragma unroll
                    for (int k = 0; k < KM; ++k) {
                        const double2 wu = make_double2(0.0, sg * dt.cx[r][k]);
                        const double2 al = dt.al[r][k];
                        const double y = dt.cy[r][k];
                        const double2 wd = make_double2(-sg * al.y * y, sg * al.x * y);
                        acc = cmad(acc, wu, dt.gu[r][k]);
                        acc = cmad(acc, wd, dt.gd[r][k]);
                    }
                    Cb[e] = acc;
                }
            }
        }
        double2 ycur[4], acur[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            ycur[r] = dt.y[r];
            acur[r] = dt.a[r];
        }
        // descriptor of the tile after next, before the prefetch is issued
        const int t_after = tn + gridDim.x;
        Tile aft = nxt;
        if (t_after < ntiles) aft = tile_of(A, t_after);
        __syncthreads();
        if (tn < ntiles) {
            load_dat(A, nxt, ix, dt);
            if (t_after < ntiles) load_idx(A, aft, ix);
        }
        // MFMA commutator of tile `cur` from LDS buffer b
        const int ar = ti * 16 + li, bc = tj * 16 + li;
        dbl4 cr = {0.0, 0.0, 0.0, 0.0}, ci = {0.0, 0.0, 0.0, 0.0};
        const double2* R = Rb;
        if (vreal) {
#pragma unroll 4
            for (int k0 = 0; k0 < NP; k0 += 4) {
                const int kk = k0 + lk;
                const double va = sVr[ar * LD + kk], vb = sVr[kk * LD + bc];
                const double2 rb = R[kk * LD + bc], ra = R[ar * LD + kk];
                cr = mfma4(va, rb.x, cr);
                ci = mfma4(va, rb.y, ci);
                cr = mfma4(ra.x, -vb, cr);
                ci = mfma4(ra.y, -vb, ci);
            }
        } else {
#pragma unroll 2
            for (int k0 = 0; k0 < NP; k0 += 4) {
                const int kk = k0 + lk;
                const double var = sVr[ar * LD + kk], vai = sVi[ar * LD + kk];
                const double vbr = sVr[kk * LD + bc], vbi = sVi[kk * LD + bc];
                const double2 rb = R[kk * LD + bc], ra = R[ar * LD + kk];
                cr = mfma4(var, rb.x, cr);
                ci = mfma4(var, rb.y, ci);
                cr = mfma4(-vai, rb.y, cr);
                ci = mfma4(vai, rb.x, ci);
                cr = mfma4(ra.x, -vbr, cr);
                ci = mfma4(ra.y, -vbr, ci);
                cr = mfma4(ra.y, vbi, cr);
                ci = mfma4(ra.x, -vbi, ci);
            }
        }
        if (live) {
            const int col = tj * 16 + li;
            const int K = cur.K, so = cur.so;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = ti * 16 + lk + 4 * r;
                if (row < N && col < N) {
                    double kr = ci[r], ki = -cr[r];   // -i * C
                    const double2 x = R[row * LD + col];
                    kr = fma(g.x, x.x, fma(-g.y, x.y, kr));
                    ki = fma(g.x, x.y, fma(g.y, x.x, ki));
                    const int jr = row - so, jc = col - so;
                    if (jr >= 0 && jr < K) {
                        const double2 v = Cb[(jr * 2 + 0) * N + col];
                        kr += v.x;
                        ki += v.y;
                    }
                    if (jc >= 0 && jc < K) {
                        const double2 v = Cb[(jc * 2 + 1) * N + row];
                        kr += v.x;
                        ki += v.y;
                    }
                    const size_t gi = (size_t)(base + row * N + col);
                    // the RK4 schedule of rk_store (7 reads + 5 writes per step)
                    const double2 yc = ycur[r];
                    if (st == 0) {
                        A.out[gi] = make_double2(fma(h2, kr, x.x), fma(h2, ki, x.y));
                    } else if (st == 1) {
                        const double2 b = make_double2((x.x + yc.x) * (1.0 / 3.0), (x.y + yc.y) * (1.0 / 3.0));
                        A.acc[gi] = make_double2(fma(hb, kr, b.x), fma(hb, ki, b.y));
                        A.out[gi] = make_double2(fma(h2, kr, yc.x), fma(h2, ki, yc.y));
                    } else if (st == 2) {
                        A.out[gi] = make_double2(fma(h, kr, yc.x), fma(h, ki, yc.y));
                    } else {
                        const double2 b = make_double2(fma(x.x, 1.0 / 3.0, acur[r].x), fma(x.y, 1.0 / 3.0, acur[r].y));
                        A.out[gi] = make_double2(fma(hb, kr, b.x), fma(hb, ki, b.y));
                    }
                }
            }
        }
        if (tn >= ntiles) break;
        cur = nxt;
        nxt = aft;
        tn = t_after;
        b ^= 1;
    }
}

__global__ void reset_state(double2* y, long long total, double2* yc, long long ctotal, int* ctr) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long e = i; e < total; e += stride) y[e] = make_double2(0.0, 0.0);
    for (long long e = i; e < ctotal; e += stride) yc[e] = make_double2(0.0, 0.0);
    if (i == 0) *ctr = 0;
}
// rho_0 = |0><0|; element (0, 0) lies in a coupled column when site 0 couples (so = 0)
__global__ void set_ground(double2* y, double2* yc, const InstDev* inst, int n_inst) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_inst) return;
    const InstDev& I = inst[i];
    y[I.state_off] = make_double2(1.0, 0.0);
    if (I.so == 0 && I.K > 0) yc[I.col_off] = make_double2(1.0, 0.0);
}

}  // namespace eh

// ================================================================ host side

namespace eh {

int ctx_init(int device, eh_ctx* ctx) {
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count == 0) return fail(EH_EHIP, "no HIP device available");
    if (device < 0 || device >= count) return fail(EH_EINVAL, "device index out of range");
    HIPCHK(hipSetDevice(device));
    hipDeviceProp_t prop;
    HIPCHK(hipGetDeviceProperties(&prop, device));
    ctx->device = device;
    std::snprintf(ctx->name, sizeof(ctx->name), "%s (%s)", prop.name, prop.gcnArchName);
    ctx->n_cu = prop.multiProcessorCount;
    ctx->lds_per_block = (int)prop.sharedMemPerBlock;
    return scratch_setup(device);   // the allocation-free entry point's array on this device
}

int ctx_mem_info(eh_ctx* ctx, int64_t* free_bytes, int64_t* total_bytes) {
    HIPCHK(hipSetDevice(ctx->device));
    size_t fr = 0, tot = 0;
    HIPCHK(hipMemGetInfo(&fr, &tot));
    if (free_bytes) *free_bytes = (int64_t)fr;
    if (total_bytes) *total_bytes = (int64_t)tot;
    return EH_OK;
}

static inline double2 d2(cplx c) { return make_double2(c.real(), c.imag()); }

int launch_cfg(const Group& g, const LaunchCfg& c, const StageArgs& base, hipStream_t s) {
    StageArgs a = base;
    if (!c.cols || a.traj) {   // this configuration (or trajectory mode) gathers columns from rin
        a.cin = nullptr;
        a.cout = nullptr;
    }
    a.stream = c.stream ? 1 : 0;
    a.tile_inst += c.tile_begin;
    a.tile_ado0 += c.tile_begin;
    a.tile_nado += c.tile_begin;
    const dim3 grid(c.tile_count), block(256);
    switch (c.nt) {
#define EH_CASE(n)                                                                                 \
    case n:                                                                                        \
        if (c.ge == 256)                                                                           \
            hipLaunchKernelGGL((stage_small<n, 256>), grid, block, c.shmem, s, a);                 \
        else if (c.ge == 1024)                                                                     \
            hipLaunchKernelGGL((stage_small<n, 1024>), grid, block, c.shmem, s, a);                \
        else if (c.late)                                                                           \
            hipLaunchKernelGGL((stage_small<n, 512, true>), grid, block, c.shmem, s, a);           \
        else                                                                                       \
            hipLaunchKernelGGL((stage_small<n, 512>), grid, block, c.shmem, s, a);                 \
        break;
        EH_CASE(1) EH_CASE(2) EH_CASE(3) EH_CASE(4) EH_CASE(5) EH_CASE(6) EH_CASE(7) EH_CASE(8)
        EH_CASE(9) EH_CASE(10) EH_CASE(11) EH_CASE(12) EH_CASE(13) EH_CASE(14) EH_CASE(15) EH_CASE(16)
#undef EH_CASE
        case -32: hipLaunchKernelGGL((stage_mfma<32, true>), grid, block, c.shmem, s, a); break;
        case -34: hipLaunchKernelGGL((stage_mfma<32, true, true>), grid, block, c.shmem, s, a); break;
        case -64: hipLaunchKernelGGL((stage_mfma<64, false>), grid, block, c.shmem, s, a); break;
        case -65: hipLaunchKernelGGL(stage_mfma64, grid, dim3(512), c.shmem, s, a); break;
        case -128:   // ge: workgroups per ADO
            hipLaunchKernelGGL(stage_mfma_wide, dim3(c.tile_count, std::max(c.ge, 1)), block, 0, s, a);
            break;
        case -67: hipLaunchKernelGGL(stage_sparse, grid, dim3(512), c.shmem, s, a); break;
        case -69:   // one kernel per RK4 stage and rows per lane (7: N <= 56, 8: N <= 64)
            hipLaunchKernelGGL(sparse_rows_kernel(a.stage, sparse_rows_of(g.maxnn), c.rows1), grid, dim3(512),
                               c.shmem, s, a);
            break;
        case -33:
            // the RK coefficients of this stage are arguments (one time step per group)
            hipLaunchKernelGGL(stage_mfma_pipe<2>, dim3(std::min(c.tile_count, g.persist)), block, c.shmem, s, a,
                               c.tile_count, g.maxkn, g.h2, g.h, g.hb[a.stage]);
            break;
        default: hipLaunchKernelGGL(stage_generic, grid, block, c.shmem, s, a); break;
    }
    HIPCHK(hipGetLastError());
    return EH_OK;
}

StageArgs stage_io(const StageArgs& sa, const BatchImpl* im, int k, int st) {
    StageArgs a = sa;
    a.step_off = k;
    a.stage = st;
    a.rin = st == 0 ? im->Y.p : (st == 2 ? im->S2.p : im->S1.p);
    a.out = st == 3 ? im->Y.p : (st == 1 ? im->S2.p : im->S1.p);
    a.cin = st == 0 ? im->Yc.p : (st == 2 ? im->S2c.p : im->S1c.p);
    a.cout = st == 3 ? im->Yc.p : (st == 1 ? im->S2c.p : im->S1c.p);
    return a;
}

StageArgs stage_args(const BatchImpl* im) {
    StageArgs sa;
    std::memset(&sa, 0, sizeof(sa));
    sa.inst = im->inst.p;
    sa.tile_inst = im->tile_inst.p;
    sa.tile_ado0 = im->tile_ado0.p;
    sa.tile_nado = im->tile_nado.p;
    sa.step_ctr = im->step_ctr.p;
    sa.vreal = im->vreal.p;
    sa.V = im->V.p;
    sa.Vpad = im->Vpad.p;
    sa.nb_up = im->nb_up.p;
    sa.nb_dn = im->nb_dn.p;
    sa.cx = im->cx.p;
    sa.cy = im->cy.p;
    sa.alpha = im->alpha.p;
    sa.alpha_t = im->alpha_t.p;
    sa.damp = im->damp.p;
    sa.sp_cnt = im->sp_cnt.p;
    sa.sp_idx = im->sp_idx.p;
    sa.gsign = 1.0;
    sa.y = im->Y.p;
    sa.acc = im->ACC.p;
    sa.vstride = (long long)(im->V.n / 2);
    sa.vpstride = (long long)(im->Vpad.n / 2);
    sa.fz = im->fz.p;
    return sa;
}

PrologueArgs prologue_args(const BatchImpl* im) {
    PrologueArgs pa;
    std::memset(&pa, 0, sizeof(pa));
    pa.inst = im->inst.p;
    pa.step_ctr = im->step_ctr.p;
    pa.H = im->H.p;
    pa.Xp = im->Xp.p;
    pa.field = im->field.p;
    pa.pidx = im->pidx.p;
    pa.R = im->R.p;
    pa.y = im->Y.p;
    pa.V = im->V.p;
    pa.Vpad = im->Vpad.p;
    pa.P = im->P.p;
    pa.readout = 1;
    return pa;
}

void launch_reset_raw(hipStream_t s, double2* y, long long total, double2* yc, long long ctotal, int* ctr,
                      const InstDev* inst, int n_inst) {
    const long long mx = std::max(total, ctotal);
    const int blocks = (int)std::min<long long>(std::max<long long>(1, (mx + 255) / 256), 65536);
    hipLaunchKernelGGL(reset_state, dim3(blocks), dim3(256), 0, s, y, total, yc, ctotal, ctr);
    hipLaunchKernelGGL(set_ground, dim3((n_inst + 255) / 256), dim3(256), 0, s, y, yc, inst, n_inst);
}

// every state to rho_0 = |0><0| (and its column copy), step counter to 0
static void launch_reset(Batch* b) {
    BatchImpl* im = b->impl;
    launch_reset_raw(im->stream, im->Y.p, (long long)b->total_elems, im->Yc.p, (long long)im->Yc.n, im->step_ctr.p,
                     im->inst.p, b->n_inst);
    PrologueArgs pa = prologue_args(im);   // V of step 0 (the fused graph's first step reads it)
    pa.readout = 0;
    hipLaunchKernelGGL(prologue, dim3(b->n_inst), dim3(256), 0, im->stream, pa);
}

// The forward graph: `chunk` steps, each group on its own branch (instances of
// different groups are independent, so their stage chains overlap), joined before
// the step-counter tick.  A step is four stage launches per group: the readouts
// and the next step's V are done inside stages 0 and 3 (fused steps), except with
// the persistent pipelined kernel, whose steps keep the prologue launch.
static int capture_forward(Batch* b) {
    BatchImpl* im = b->impl;
    if (im->exec) (void)hipGraphExecDestroy(im->exec);
    if (im->graph) (void)hipGraphDestroy(im->graph);
    im->exec = nullptr;
    im->graph = nullptr;
    StageArgs sa = stage_args(im);
    const PrologueArgs pa = prologue_args(im);
    bool fused = true;
    for (const Group& g : im->groups) fused = fused && g.cand[g.sel].nt != -33;
    sa.fused = fused ? 1 : 0;
    HIPCHK(hipStreamBeginCapture(im->stream, hipStreamCaptureModeThreadLocal));
    int rc = EH_OK;
    hipError_t e = hipEventRecord(im->fork_ev, im->stream);
    for (size_t gi = 0; gi < im->groups.size() && rc == EH_OK && e == hipSuccess; gi++) {
        const Group& g = im->groups[gi];
        hipStream_t s = gi == 0 ? im->stream : im->side[gi - 1];
        if (gi) e = hipStreamWaitEvent(s, im->fork_ev, 0);
        PrologueArgs pg = pa;
        pg.ilist = im->ginst.p + g.inst_begin;
        for (int k = 0; k < b->chunk && rc == EH_OK && e == hipSuccess; k++) {
            pg.step_off = k;
            if (!fused) hipLaunchKernelGGL(prologue, dim3(g.inst_count), dim3(256), 0, s, pg);
            for (int st = 0; st < 4 && rc == EH_OK; st++) rc = launch_stage(g, stage_io(sa, im, k, st), s);
        }
        if (gi && e == hipSuccess) {
            e = hipEventRecord(im->join_ev[gi - 1], s);
            if (e == hipSuccess) e = hipStreamWaitEvent(im->stream, im->join_ev[gi - 1], 0);
        }
    }
    hipLaunchKernelGGL(tick, dim3(1), dim3(1), 0, im->stream, im->step_ctr.p, b->chunk);
    hipGraph_t gr = nullptr;
    const hipError_t ee = hipStreamEndCapture(im->stream, &gr);
    if (rc || e != hipSuccess || ee != hipSuccess) {
        if (ee == hipSuccess && gr) (void)hipGraphDestroy(gr);
        if (rc) return rc;
        return fail(EH_EHIP, std::string("forward graph capture: ") + hipGetErrorString(e != hipSuccess ? e : ee));
    }
    im->graph = gr;
    HIPCHK(hipGraphInstantiate(&im->exec, im->graph, nullptr, nullptr, 0));
    return EH_OK;
}

// The host side of batch_create: every check, table, tile list, group and launch
// candidate of the n instances, with no device allocation or copy (the only device
// call is hipFuncSetAttribute for the candidates whose LDS exceeds 64 KB).  Fills
// b's sizes and per-instance info.
int build_host_batch(eh_ctx* ctx, int n, const eh_instance* in, Batch* b, HostBatch& hb) {
    b->ctx = ctx;
    b->n_inst = n;
    b->info.assign((size_t)n, Batch::Info());
    b->ado_updates = 0;
    b->max_steps = 0;
    std::vector<InstDev>& hinst = hb.inst;
    hinst.assign((size_t)n, InstDev());
    std::vector<int>&t_inst = hb.t_inst, &t_ado0 = hb.t_ado0, &t_nado = hb.t_nado;
    std::vector<int>&hup = hb.up, &hdn = hb.dn, &hpidx = hb.pidx;
    std::vector<uint8_t>& hado = hb.ado;
    std::vector<double>&hcx = hb.cx, &hcy = hb.cy;
    // Hierarchy tables (neighbour indices, occupations, link coefficients) once per
    // distinct (d, L, alpha): they depend on nothing else, so a grid of propagations
    // with one bath shares a few small, cache-resident tables.
    std::map<std::vector<double>, int64_t> hier_off;
    std::map<std::vector<short>, int> sp_rows;   // distinct sparsity patterns of V -> first row
    std::vector<int> isp((size_t)n, 0);          // largest pattern row count per instance
    hb.spcnt.clear();
    hb.spidx.clear();
    std::vector<uint8_t>& hvreal = hb.vreal;
    std::vector<double2>&hH = hb.H, &hXp = hb.Xp, &hfield = hb.field, &hR = hb.R, &hal = hb.al, &halt = hb.alt,
                         &hdamp = hb.damp;
    std::vector<long long>& hado0 = hb.ado0;
    int64_t state_off = 0, p_off = 0, lk_off = 0, col_off = 0, vp_off = 0;
    int ado_off = 0, v_off = 0, r_off = 0, step_off = 0, al_off = 0;
    // instances by N class: 1..16 (each N its own class), 17 (16 < N <= 32), 18 (32 < N <= 64),
    // 19 (N > 64)
    std::vector<std::vector<int>> by_class(20);
    auto env_on = [](const char* v) { const char* e = std::getenv(v); return e && e[0] == '1'; };
    const bool use_mfma = !env_on("ETOHEOM_NO_MFMA");
    const bool use_pipe = env_on("ETOHEOM_PIPE");   // opt-in default: slower than stage_mfma<32> as compiled

    for (int i = 0; i < n; i++) {
        const eh_instance& q = in[i];
        InstPlan ip;   // the host-side checks (sizes, 32-bit in-instance offsets, time grid)
        if (const int rc0 = check_instance(q, ip)) return rc0;
        const eh_system& sys = *q.sys;
        const int N = sys.N;
        Hierarchy h;
        int rc = build_hierarchy(sys.K * sys.Km, sys.L, h);
        if (rc) return rc;
        Tables tb;
        rc = build_tables(h, sys, tb);
        if (rc) return rc;
        Schedule s;
        rc = build_schedule(sys, q.mode, q.nv, s);
        if (rc) return rc;
        const int NN = N * N;
        const int M = (int)h.M;
        // readout operators
        int nread = 0;
        if (q.mode == EH_MODE_2DES) {
            nread = 1;
            for (int e = 0; e < NN; e++) hR.push_back(make_double2(q.X4[2 * (3 * NN + e)], q.X4[2 * (3 * NN + e) + 1]));
        } else {
            if (q.n_sites > N) return fail(EH_ERANGE, "number of output sites exceeds the system size");
            nread = q.n_sites;
            for (int r = 0; r < q.n_sites; r++) {
                const int st = q.sites[r];
                if (st < 1 || st > N) return fail(EH_ERANGE, "site " + std::to_string(st) + " is out of range");
                for (int e = 0; e < NN; e++) hR.push_back(make_double2(e == (st - 1) * N + (st - 1) ? 1.0 : 0.0, 0.0));
            }
        }
        bool real_h = true, sym_h = true;
        for (int e = 0; e < NN; e++) {
            hH.push_back(make_double2(q.H[2 * e], q.H[2 * e + 1]));
            if (q.H[2 * e + 1] != 0.0) real_h = false;
            const int ea = e / N, eb = e - ea * N;
            if (q.H[2 * e] != q.H[2 * (eb * N + ea)]) sym_h = false;
        }
        for (int qq = 0; qq < 3; qq++)
            for (int a = 0; a < N; a++)
                for (int bb = 0; bb < N; bb++) {
                    if (q.mode != EH_MODE_2DES) {
                        hXp.push_back(make_double2(0.0, 0.0));
                        continue;
                    }
                    const int e = qq == 0 ? a * N + bb : bb * N + a;   // X[1], X[2] transposed
                    hXp.push_back(make_double2(q.X4[2 * (qq * NN + e)], q.X4[2 * (qq * NN + e) + 1]));
                }
        // the sparsity pattern of V(t) = H - sum_q e_q(t) X'_q over the whole run, made
        // symmetric; instances with the same pattern share its rows
        int sp_row = -1, sp_max = 0;
        {
            const double2* Hh = hH.data() + (size_t)v_off;
            const double2* Xh = hXp.data() + (size_t)3 * v_off;
            auto z = [](double2 v) { return v.x != 0.0 || v.y != 0.0; };
            std::vector<short> pat((size_t)N * kSpCap, (short)-1);
            std::vector<int> cnt((size_t)N, 0);
            for (int a = 0; a < N; a++) {
                for (int c = 0; c < N; c++) {
                    bool on = z(Hh[a * N + c]) || z(Hh[c * N + a]);
                    for (int qq = 0; qq < 3 && !on; qq++) on = z(Xh[qq * NN + a * N + c]) || z(Xh[qq * NN + c * N + a]);
                    if (!on) continue;
                    if (cnt[(size_t)a] < kSpCap) pat[(size_t)a * kSpCap + cnt[(size_t)a]] = (short)c;
                    cnt[(size_t)a]++;
                }
                sp_max = std::max(sp_max, cnt[(size_t)a]);
            }
            if (sp_max <= kSpCap) {
                const auto it = sp_rows.find(pat);
                if (it != sp_rows.end()) {
                    sp_row = it->second;
                } else {
                    sp_row = (int)hb.spcnt.size();
                    sp_rows[pat] = sp_row;
                    hb.spcnt.insert(hb.spcnt.end(), cnt.begin(), cnt.end());
                    hb.spidx.insert(hb.spidx.end(), pat.begin(), pat.end());
                }
            }
        }
        isp[(size_t)i] = sp_max;
        for (int st = 0; st < s.nsteps; st++) {
            for (int qq = 0; qq < 3; qq++) hfield.push_back(d2(s.e[(size_t)3 * st + qq]));
            hpidx.push_back(s.pidx[(size_t)st]);
            hvreal.push_back((real_h && !s.active[(size_t)st]) ? 1 : 0);
        }
        for (int m = 0; m < M; m++) hdamp.push_back(d2(tb.damp[(size_t)m]));
        int64_t inst_lk;
        {
            std::vector<double> key = {(double)h.d, (double)sys.L, (double)sys.K, (double)sys.Km};
            key.insert(key.end(), sys.alpha, sys.alpha + 2 * sys.K * sys.Km);
            const auto it = hier_off.find(key);
            if (it != hier_off.end()) {
                inst_lk = it->second;
            } else {
                inst_lk = lk_off;
                hier_off[key] = lk_off;
                hup.insert(hup.end(), tb.up.begin(), tb.up.end());
                hado.insert(hado.end(), h.ado.begin(), h.ado.end());
                hdn.insert(hdn.end(), tb.dn.begin(), tb.dn.end());
                hcx.insert(hcx.end(), tb.cx.begin(), tb.cx.end());
                hcy.insert(hcy.end(), tb.cy.begin(), tb.cy.end());
                lk_off += (int64_t)M * h.d;
            }
        }
        for (int e = 0; e < sys.K * sys.Km; e++) {
            hal.push_back(make_double2(sys.alpha[2 * e], sys.alpha[2 * e + 1]));
            halt.push_back(make_double2(sys.alpha_t[2 * e], sys.alpha_t[2 * e + 1]));
        }
        InstDev& d = hinst[i];
        d.state_off = state_off;
        d.p_off = p_off;
        d.N = N;
        d.M = M;
        d.ado_off = ado_off;
        d.nsteps = s.nsteps;
        d.v_off = v_off;
        d.step_off = step_off;
        d.nread = nread;
        d.npoints = s.npoints;
        d.r_off = r_off;
        d.lk_off = inst_lk;
        d.d = h.d;
        d.K = sys.K;
        d.Km = sys.Km;
        d.so = sys.site_offset;
        d.al_off = al_off;
        d.col_off = col_off;
        col_off += (int64_t)M * sys.K * N;
        d.hsym = real_h && sym_h ? 1 : 0;
        d.sp_off = sp_row;
        d.vp_off = -1;
        if (N > 32 && N <= 64) {
            d.vp_off = vp_off;
            vp_off += 2 * 64 * 64;
        }
        d.h2 = s.h2;
        d.h = s.h;
        for (int k = 0; k < 4; k++) d.hb[k] = s.hb[k];
        hado0.push_back(state_off);
        Batch::Info& inf = b->info[(size_t)i];
        inf.N = N;
        inf.M = M;
        inf.nsteps = s.nsteps;
        inf.npoints = s.npoints;
        inf.nread = nread;
        inf.mode = q.mode;
        inf.state_off = state_off;
        inf.p_off = p_off;
        b->ado_updates += (int64_t)M * s.nsteps;
        b->max_steps = std::max(b->max_steps, s.nsteps);
        by_class[(size_t)(N <= 16 ? N : (N <= 32 ? 17 : (N <= 64 ? 18 : 19)))].push_back(i);
        state_off += (int64_t)M * NN;
        p_off += (int64_t)nread * s.npoints;
        ado_off += M;
        v_off += NN;
        r_off += nread * NN;
        step_off += s.nsteps;
        al_off += sys.K * sys.Km;
    }
    b->total_ados = ado_off;
    b->total_elems = state_off;
    hb.state_elems = state_off;
    hb.p_elems = p_off;
    hb.col_elems = col_off;
    hb.vp_elems = vp_off;
    // Groups and their launch candidates.  Tiles are (instance, first ADO, ADO count)
    // triples; a group's candidates with equal ADOs per tile share one tile range.
    std::vector<int>& hginst = hb.ginst;
    hb.groups.clear();
    // XCD-aware tile order: workgroups go to the XCDs round-robin (workgroup w to XCD
    // w mod xcd), each XCD with its own L2, so a range of tiles is laid out such that
    // XCD x runs the x-th contiguous slice of the (instance, ADO) order: an instance's
    // neighbour gathers then stay in one L2 instead of all of them.
    const int xcd = env_on("ETOHEOM_NO_XCD") ? 1 : (ctx->n_cu % 8 == 0 && ctx->n_cu >= 64 ? 8 : 1);
    auto xcd_order = [&](int begin, int count) {
        if (count < 2 * xcd) return;
        std::vector<int> qi(t_inst.begin() + begin, t_inst.begin() + begin + count);
        std::vector<int> qa(t_ado0.begin() + begin, t_ado0.begin() + begin + count);
        std::vector<int> qn(t_nado.begin() + begin, t_nado.begin() + begin + count);
        std::vector<int> head(xcd), end(xcd);
        for (int x = 0; x < xcd; x++) {
            head[x] = (int)((long long)count * x / xcd);
            end[x] = (int)((long long)count * (x + 1) / xcd);
        }
        for (int p = 0; p < count; p++) {
            int x = p % xcd;
            if (head[x] == end[x]) {   // this slice is done: take from the fullest one
                int best = 0;
                for (int y = 1; y < xcd; y++)
                    if (end[y] - head[y] > end[best] - head[best]) best = y;
                x = best;
            }
            const int src = head[x]++;
            t_inst[(size_t)(begin + p)] = qi[(size_t)src];
            t_ado0[(size_t)(begin + p)] = qa[(size_t)src];
            t_nado[(size_t)(begin + p)] = qn[(size_t)src];
        }
    };
    for (int cls = 1; cls <= 19; cls++) {
        const auto& members = by_class[(size_t)cls];
        if (members.empty()) continue;
        Group g;
        g.inst_begin = (int)hginst.size();
        g.inst_count = (int)members.size();
        hginst.insert(hginst.end(), members.begin(), members.end());
        g.persist = 2 * ctx->n_cu;
        int maxnn = 0, maxkn = 0, maxd = 0, maxsp = 0;
        bool pipe_ok = cls == 17;
        bool tp_ok = true;   // two-phase gathers: 2 K N <= 512 and K_m <= 2 (and a bath) everywhere
        const InstDev& f0 = hinst[(size_t)members[0]];
        for (int i : members) {
            const InstDev& di = hinst[(size_t)i];
            maxnn = std::max(maxnn, di.N * di.N);
            maxkn = std::max(maxkn, 2 * di.K * di.N);
            maxd = std::max(maxd, di.d);
            maxsp = std::max(maxsp, isp[(size_t)i]);
            // the pipelined kernel takes the RK coefficients as arguments: one time step per group
            if (di.Km > 2 || 2 * di.K * di.N > 512 || di.h != f0.h) pipe_ok = false;
            if (di.Km > 2 || di.Km < 1 || di.K < 1 || 2 * di.K * di.N > 512) tp_ok = false;
        }
        g.maxnn = maxnn;
        g.maxkn = maxkn;
        g.maxd = maxd;
        g.maxsp = maxsp;
        // the sparse-commutator kernel: every instance's pattern within its row capacity
        const bool sparse_ok = maxsp <= kSpCap && !env_on("ETOHEOM_NO_SPARSE");
        // Non-temporal state stores in the small-N kernels by default when the group's four
        // state buffers exceed 512 MB, twice the last-level cache (measured on the 2D grid:
        // the 1.6 GB N = 11 group's stage 3 % faster with them, the 224 MB N = 7 group's
        // 12 % slower; the MFMA kernels: the 165 MB heptamer 19 % slower, the 355 MB chain
        // unchanged, so they never stream); the autotuner tries the other.
        double gbytes = 0.0;
        for (int i : members) gbytes += 4.0 * sizeof(double2) * hinst[(size_t)i].M * hinst[(size_t)i].N * hinst[(size_t)i].N;
        const bool stream = gbytes > 512.0 * (1 << 20);
        g.h = f0.h;
        g.h2 = f0.h2;
        for (int k = 0; k < 4; k++) g.hb[k] = f0.hb[k];
        std::vector<std::pair<int, std::pair<int, int>>> ranges;   // ADOs per tile -> (begin, count)
        // with the XCD order below, the instances go to the slices round-robin (member k to
        // slice k mod xcd): neighbouring grid points, which have their pulses on at the same
        // steps, then load every XCD alike instead of a few (measured on 256 propagations of
        // the 10-site chain, a 64 t1 x 4 t2 grid: RK4 steps 250-1249 in 50-step replays
        // 3.35-3.57 s -> 3.29-3.36 s, where the steps without any pulse take 3.25 s)
        std::vector<int> order(members.begin(), members.end());
        if (xcd > 1 && cls > 16) {
            order.clear();
            for (int x = 0; x < xcd; x++)
                for (size_t k = (size_t)x; k < members.size(); k += (size_t)xcd) order.push_back(members[k]);
        }
        auto tiles = [&](int G) {
            for (const auto& r : ranges)
                if (r.first == G) return r.second;
            const int begin = (int)t_inst.size();
            for (int i : order) {
                const int M = hinst[(size_t)i].M;
                for (int a0 = 0; a0 < M; a0 += G) {
                    t_inst.push_back(i);
                    t_ado0.push_back(a0);
                    t_nado.push_back(std::min(G, M - a0));
                }
            }
            const std::pair<int, int> r(begin, (int)t_inst.size() - begin);
            // measured: helps the N > 16 kernels; the small-N forward and the adjoint
            // (atomics of one instance from one XCD) are faster without it
            if (xcd > 1 && cls > 16) xcd_order(begin, r.second);
            ranges.push_back({G, r});
            return r;
        };
        const size_t cpb = (size_t)maxkn * sizeof(double2);   // coupling rows/columns
        // late < 0: the deferred-load small-N variant (GE = 512) when the group has at least
        // two workgroups per CU (measured: it pays where occupancy matters, not on a few tiles)
        auto add = [&](int nt, int ge, bool cols, bool stream, int late = 0) {
            LaunchCfg c;
            c.nt = nt;
            c.ge = ge;
            c.cols = cols;
            c.stream = stream;
            const int G = nt > 0 ? (maxnn >= ge ? 1 : ge / maxnn) : 1;
            const std::pair<int, int> r = tiles(G);
            c.tile_begin = r.first;
            c.tile_count = r.second;
            c.late = nt > 0 && ge == 512 && (late > 0 || (late < 0 && c.tile_count >= 2 * ctx->n_cu));
            if (nt == -33)
                c.shmem = (size_t)2 * 32 * 33 * sizeof(double2) + (size_t)2 * 32 * 33 * sizeof(double) + 2 * cpb;
            else if (nt == -32 || nt == -34)
                c.shmem = (size_t)2 * 32 * 33 * sizeof(double2) + cpb;
            else if (nt == -64)
                c.shmem = (size_t)64 * 65 * sizeof(double2) + cpb;
            else if (nt == -65) {
                int mn = 1;
                while (mn * mn < maxnn) mn++;
                c.shmem = (size_t)mn * (mn + 1) * sizeof(double2) + cpb;
            }
            else if (nt == -128)
                c.shmem = 0;   // stage_mfma_wide: nothing in LDS
            else if (nt == -67 || nt == -69) {   // stage_sparse: rho, the couplings, the column lists
                int mn = 1;
                while (mn * mn < maxnn) mn++;
                const size_t ps = kSpCap + 1;
                c.shmem = (size_t)mn * (mn + 1) * sizeof(double2) + cpb + (size_t)mn * ps * sizeof(double) +
                          (((size_t)mn * ps * sizeof(short) + 3) & ~(size_t)3) + (size_t)mn * sizeof(int);
                // stage 1 of stage_sparse_rows exchanges half of rho V's rows at a time past rho
                // (and holds one array more: seven rows per lane, N <= 56, fit the registers)
                c.rows1 = nt == -69 && mn <= 56 &&
                          (size_t)mn * (mn + 1) * sizeof(double2) +
                                  (size_t)((mn + 1) / 2) * (mn | 1) * sizeof(double2) <= c.shmem;
            }
            else if (nt > 0)
                c.shmem = (size_t)G * cpb;   // the coupling rows/columns of G ADOs (the rest is static)
            else
                c.shmem = (size_t)maxnn * sizeof(double2) + cpb;
            if ((int)c.shmem > ctx->lds_per_block) return;   // not a candidate on this device
            if (c.shmem > 65536) {
                const void* fn = nt == -64   ? (const void*)stage_mfma<64, false>
                                 : nt == -65 ? (const void*)stage_mfma64
                                 : nt == -67 ? (const void*)stage_sparse
                                 : nt == -69 ? (const void*)sparse_rows_kernel(0, sparse_rows_of(maxnn), false)
                                 : (nt == -33 ? (const void*)stage_mfma_pipe<2> : (const void*)stage_generic);
                if (hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)c.shmem) != hipSuccess)
                    return;
                for (int st = 0; nt == -69 && st < 4; st++)   // one kernel per RK4 stage
                    if (hipFuncSetAttribute((const void*)sparse_rows_kernel(st, sparse_rows_of(maxnn), c.rows1),
                                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)c.shmem) != hipSuccess)
                        return;
            }
            g.cand.push_back(c);
        };
        // column copies: measured to pay off for N > 16 (long strided columns), not below,
        // where the small-N kernel has no column-copy path
        if (cls <= 16) {
            add(cls, 512, false, stream, -1);   // the base configuration: the adjoint's small kernels tile as it does
            add(cls, 256, false, stream);
            add(cls, 1024, false, stream);
            add(0, 0, false, stream);
        } else if (cls == 17) {   // base: two-phase gathers where the instances allow
            if (use_mfma && tp_ok) add(-34, 0, false, stream);
            if (use_mfma && tp_ok) add(-34, 0, true, stream);
            if (use_mfma) add(-32, 0, true, stream);
            add(0, 0, true, stream);
            if (use_mfma) add(-32, 0, false, stream);
            if (use_mfma && pipe_ok) add(-33, 0, false, stream);
        } else if (cls == 19) {   // N > 64: operands from global memory; LDS-staged VALU where it fits
            // the wide kernel on one workgroup per ADO, and with the ADO's 16 x 16 tiles spread
            // over as many workgroups as give each wave at most four
            int mn = 1;
            while (mn * mn < maxnn) mn++;
            const int nt16 = (mn + 15) / 16, split = (nt16 * nt16 + 15) / 16;
            if (use_mfma) add(-128, 1, false, stream);
            if (use_mfma && split > 1) add(-128, split, false, stream);
            add(0, 0, false, stream);
        } else {   // 32 < N <= 64: stage_mfma64 gathers columns directly (measured faster at N = 56)
            if (use_mfma) add(-65, 0, false, stream);
            add(0, 0, true, stream);
            if (sparse_ok) add(-67, 0, false, stream);
            if (sparse_ok) add(-69, 0, false, stream);   // the default (below)
            if (use_mfma) {
                add(-65, 0, true, stream);
                add(-64, 0, true, stream);
            }
        }
        // the small-N base configuration with the other store hint (the autotuner's choice)
        if (cls <= 16 && !g.cand.empty()) {
            const LaunchCfg b0 = g.cand[0];
            add(b0.nt, b0.ge, b0.cols, !stream, b0.late ? 1 : 0);
            add(b0.nt, b0.ge, b0.cols, stream, b0.late ? 0 : 1);   // and with the other load placement
        }
        if (g.cand.empty() || (cls <= 16 && g.cand[0].nt != cls))
            return fail(EH_ESHAPE, "state tile exceeds LDS");
        for (size_t c = 0; c < g.cand.size(); c++) {
            if (use_pipe && g.cand[c].nt == -33) g.sel = (int)c;
            if (g.cand[c].nt == -69) g.sel = (int)c;   // measured (256 x N = 56 chain): 78.2 -> 73.9 ms per RK4 step
            // the wide kernel over several workgroups per ADO where the ADOs alone leave the chip
            // short of work (measured per RK4 step: N = 137 x 561 ADOs 3.66 -> 2.80 ms, N = 256 x
            // 7 ADOs 6.84 -> 0.61 ms; N = 67 x 2300 ADOs 1.90 one per ADO, 1.92 two)
            if (g.cand[c].nt == -128 && g.cand[c].ge > 1 && g.cand[c].tile_count < 8 * ctx->n_cu) g.sel = (int)c;
        }
        g.nt = g.cand[0].nt;
        g.tile_begin = g.cand[0].tile_begin;
        g.tile_count = g.cand[0].tile_count;
        g.shmem = g.cand[0].shmem;
        hb.groups.push_back(g);
    }
    return EH_OK;
}

int batch_create(eh_ctx* ctx, int n, const eh_instance* in, Batch** out) {
    if (!ctx || !in || n <= 0 || !out) return fail(EH_EINVAL, "batch_create: bad arguments");
    HIPCHK(hipSetDevice(ctx->device));
    std::unique_ptr<Batch> b(new Batch());
    std::unique_ptr<BatchImpl> im(new BatchImpl());
    HostBatch hb;
    int rc = build_host_batch(ctx, n, in, b.get(), hb);
    if (rc) return rc;
    im->groups = hb.groups;
    const int64_t state_off = hb.state_elems, p_off = hb.p_elems, col_off = hb.col_elems, vp_off = hb.vp_elems;
    if ((rc = im->inst.upload(hb.inst)) || (rc = im->tile_inst.upload(hb.t_inst)) ||
        (rc = im->tile_ado0.upload(hb.t_ado0)) || (rc = im->tile_nado.upload(hb.t_nado)) ||
        (rc = im->nb_up.upload(hb.up)) || (rc = im->nb_dn.upload(hb.dn)) || (rc = im->cx.upload(hb.cx)) ||
        (rc = im->cy.upload(hb.cy)) || (rc = im->pidx.upload(hb.pidx)) || (rc = im->vreal.upload(hb.vreal)) ||
        (rc = im->H.upload(hb.H)) || (rc = im->Xp.upload(hb.Xp)) || (rc = im->field.upload(hb.field)) ||
        (rc = im->R.upload(hb.R)) || (rc = im->alpha.upload(hb.al)) || (rc = im->alpha_t.upload(hb.alt)) ||
        (rc = im->damp.upload(hb.damp)) || (rc = im->ado0_off.upload(hb.ado0)) || (rc = im->ginst.upload(hb.ginst)) ||
        (rc = im->sp_cnt.upload(hb.spcnt)) || (rc = im->sp_idx.upload(hb.spidx)))
        return rc;
    if ((rc = im->step_ctr.alloc(1)) || (rc = im->V.alloc(2 * hb.H.size())) || (rc = im->P.alloc((size_t)p_off)) ||
        (rc = im->Y.alloc((size_t)state_off)) || (rc = im->ACC.alloc((size_t)state_off)) ||
        (rc = im->S1.alloc((size_t)state_off)) || (rc = im->S2.alloc((size_t)state_off)) ||
        (rc = im->Yc.alloc((size_t)col_off)) || (rc = im->S1c.alloc((size_t)col_off)) ||
        (rc = im->S2c.alloc((size_t)col_off)) || (rc = im->Vpad.alloc(2 * (size_t)vp_off)))
        return rc;
    FusedDev fz;
    fz.V = im->V.p;
    fz.Vpad = im->Vpad.p;
    fz.vstride = (long long)(im->V.n / 2);
    fz.vpstride = (long long)(im->Vpad.n / 2);
    fz.H = im->H.p;
    fz.Xp = im->Xp.p;
    fz.field = im->field.p;
    fz.pidx = im->pidx.p;
    fz.R = im->R.p;
    fz.P = im->P.p;
    if ((rc = im->fz.upload(std::vector<FusedDev>{fz}))) return rc;
    im->hinst = std::move(hb.inst);
    im->hup = std::move(hb.up);
    im->hdn = std::move(hb.dn);
    im->hcx = std::move(hb.cx);
    im->hcy = std::move(hb.cy);
    im->hal = std::move(hb.al);
    im->halt = std::move(hb.alt);
    im->hdamp = std::move(hb.damp);
    im->hado = std::move(hb.ado);
    im->htile_inst = std::move(hb.t_inst);
    HIPCHK(hipStreamCreateWithFlags(&im->stream, hipStreamNonBlocking));
    HIPCHK(hipEventCreate(&im->ev0));
    HIPCHK(hipEventCreate(&im->ev1));

    HIPCHK(hipEventCreateWithFlags(&im->fork_ev, hipEventDisableTiming));
    for (size_t gi = 1; gi < im->groups.size(); gi++) {
        hipStream_t q = nullptr;
        HIPCHK(hipStreamCreateWithFlags(&q, hipStreamNonBlocking));
        im->side.push_back(q);
        hipEvent_t e = nullptr;
        HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        im->join_ev.push_back(e);
    }
    // graph of `chunk` steps
    const int per_step = 4 * (int)im->groups.size();
    b->chunk = std::max(1, std::min(b->max_steps, 256 / per_step));
    im->nchunks = (b->max_steps + b->chunk - 1) / b->chunk;
    b->impl = im.get();
    if ((rc = capture_forward(b.get()))) {
        b->impl = nullptr;
        return rc;
    }
    im.release();
    *out = b.release();
    return EH_OK;
}

void batch_destroy(Batch* b) {
    if (!b) return;
    if (b->ctx) (void)hipSetDevice(b->ctx->device);
    if (b->impl) {
        if (b->impl->stream) (void)hipStreamSynchronize(b->impl->stream);
        delete b->impl;
    }
    delete b;
}

int batch_run(Batch* b) {
    if (!b || !b->impl) return fail(EH_EINVAL, "batch_run: null batch");
    BatchImpl* im = b->impl;
    HIPCHK(hipSetDevice(b->ctx->device));
    launch_reset(b);
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(im->ev0, im->stream));
    if (im->adj) {
        const int rc = adj_forward(b);
        if (rc) return rc;
    } else {
        for (int c = 0; c < im->nchunks; c++) HIPCHK(hipGraphLaunch(im->exec, im->stream));
    }
    HIPCHK(hipEventRecord(im->ev1, im->stream));
    im->ran = true;
    im->started = false;
    b->synced = false;
    return EH_OK;
}

// The forward graph with `chunk` steps per replay (re-captured).
int batch_set_chunk(Batch* b, int chunk) {
    if (!b || !b->impl) return fail(EH_EINVAL, "batch_set_chunk: null batch");
    BatchImpl* im = b->impl;
    if (im->adj) return fail(EH_EINVAL, "batch_set_chunk: set the chunk before enabling the adjoint");
    if (chunk <= 0 || chunk > std::max(1, b->max_steps)) return fail(EH_ERANGE, "batch_set_chunk: chunk out of range");
    HIPCHK(hipSetDevice(b->ctx->device));
    HIPCHK(hipStreamSynchronize(im->stream));
    b->chunk = chunk;
    im->nchunks = (b->max_steps + chunk - 1) / chunk;
    im->started = false;
    return capture_forward(b);
}

// A propagation in pieces: eh_batch_start resets every state; each eh_batch_advance
// then replays the forward graph for the next nsteps steps (a multiple of the chunk),
// between HIP events on the batch stream.  Steps past an instance's last do nothing.
int batch_start(Batch* b) {
    if (!b || !b->impl) return fail(EH_EINVAL, "batch_start: null batch");
    BatchImpl* im = b->impl;
    if (im->adj) return fail(EH_EINVAL, "batch_start: the adjoint needs the checkpointed run (eh_batch_run)");
    HIPCHK(hipSetDevice(b->ctx->device));
    launch_reset(b);
    HIPCHK(hipGetLastError());
    im->started = true;
    im->pos = 0;
    b->synced = false;
    return EH_OK;
}

int batch_advance(Batch* b, int nsteps) {
    if (!b || !b->impl) return fail(EH_EINVAL, "batch_advance: null batch");
    BatchImpl* im = b->impl;
    if (!im->started) return fail(EH_EINVAL, "batch_advance: call eh_batch_start first");
    if (nsteps <= 0 || nsteps % b->chunk) return fail(EH_EINVAL, "batch_advance: steps must be a positive multiple of the chunk");
    HIPCHK(hipSetDevice(b->ctx->device));
    HIPCHK(hipEventRecord(im->ev0, im->stream));
    for (int c = 0; c < nsteps / b->chunk; c++) HIPCHK(hipGraphLaunch(im->exec, im->stream));
    HIPCHK(hipEventRecord(im->ev1, im->stream));
    im->pos += nsteps;
    im->ran = true;
    b->synced = false;
    return EH_OK;
}

int batch_sync(Batch* b) {
    if (!b || !b->impl) return fail(EH_EINVAL, "batch_sync: null batch");
    BatchImpl* im = b->impl;
    if (!im->ran) return fail(EH_EINVAL, "batch_sync: the batch has not run (its readouts are undefined)");
    HIPCHK(hipSetDevice(b->ctx->device));
    HIPCHK(hipStreamSynchronize(im->stream));
    b->P_host.resize((size_t)im->P.n);
    if (im->P.n) HIPCHK(hipMemcpy(b->P_host.data(), im->P.p, im->P.n * sizeof(double2), hipMemcpyDeviceToHost));
    b->synced = true;
    return EH_OK;
}

int batch_elapsed_ms(Batch* b, double* ms) {
    if (!b || !b->impl || !b->impl->ran) return fail(EH_EINVAL, "batch has not run");
    float f = 0;
    HIPCHK(hipEventElapsedTime(&f, b->impl->ev0, b->impl->ev1));
    *ms = f;
    return EH_OK;
}

// Per-kernel timing pass: `nsteps` steps launched one kernel at a time (no graph)
// with HIP events around every launch on the batch stream.  out[0] = mean
// prologue duration, out[1..4] = mean duration of stage 0..3 (all groups of a
// stage together), in ms.  Leaves the state advanced by nsteps steps.
int batch_profile(Batch* b, int nsteps, double* out) {
    if (!b || !b->impl || nsteps <= 0 || !out) return fail(EH_EINVAL, "batch_profile: bad arguments");
    BatchImpl* im = b->impl;
    HIPCHK(hipSetDevice(b->ctx->device));
    launch_reset(b);
    im->started = false;
    std::vector<hipEvent_t> ev((size_t)6 * nsteps);
    for (auto& e : ev) HIPCHK(hipEventCreate(&e));
    const StageArgs sa = stage_args(im);
    PrologueArgs pa = prologue_args(im);
    pa.step_off = 0;
    int rc = EH_OK;
    for (int k = 0; k < nsteps && rc == EH_OK; k++) {
        hipEvent_t* e = ev.data() + (size_t)6 * k;
        (void)hipEventRecord(e[0], im->stream);
        hipLaunchKernelGGL(prologue, dim3(b->n_inst), dim3(256), 0, im->stream, pa);
        (void)hipEventRecord(e[1], im->stream);
        for (int st = 0; st < 4 && rc == EH_OK; st++) {
            for (const Group& g : im->groups)
                if ((rc = launch_stage(g, stage_io(sa, im, 0, st), im->stream))) break;
            (void)hipEventRecord(e[2 + st], im->stream);
        }
        hipLaunchKernelGGL(tick, dim3(1), dim3(1), 0, im->stream, im->step_ctr.p, 1);
    }
    hipError_t se = hipStreamSynchronize(im->stream);
    for (int q = 0; q < 5; q++) out[q] = 0.0;
    if (rc == EH_OK && se == hipSuccess) {
        for (int k = 0; k < nsteps; k++) {
            hipEvent_t* e = ev.data() + (size_t)6 * k;
            for (int q = 0; q < 5; q++) {
                float ms = 0;
                (void)hipEventElapsedTime(&ms, e[q], e[q + 1]);
                out[q] += ms / nsteps;
            }
        }
    }
    for (auto& e : ev) (void)hipEventDestroy(e);
    if (rc) return rc;
    HIPCHK(se);
    b->synced = false;
    return EH_OK;
}

// The plain per-launch path: the same kernels launched one by one from the host,
// step after step (prologue, four stages per group, counter tick), no graph.
int batch_run_eager(Batch* b) {
    if (!b || !b->impl) return fail(EH_EINVAL, "batch_run_eager: null batch");
    BatchImpl* im = b->impl;
    if (im->adj) return fail(EH_EINVAL, "batch_run_eager: the adjoint needs the checkpointed run (eh_batch_run)");
    HIPCHK(hipSetDevice(b->ctx->device));
    launch_reset(b);
    HIPCHK(hipEventRecord(im->ev0, im->stream));
    const StageArgs sa = stage_args(im);
    const PrologueArgs pa = prologue_args(im);
    for (int k = 0; k < b->max_steps; k++) {
        hipLaunchKernelGGL(prologue, dim3(b->n_inst), dim3(256), 0, im->stream, pa);
        for (int st = 0; st < 4; st++)
            for (const Group& g : im->groups) {
                const int rc = launch_stage(g, stage_io(sa, im, 0, st), im->stream);
                if (rc) return rc;
            }
        hipLaunchKernelGGL(tick, dim3(1), dim3(1), 0, im->stream, im->step_ctr.p, 1);
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(im->ev1, im->stream));
    im->ran = true;
    im->started = false;
    b->synced = false;
    return EH_OK;
}

// Autotuner: for every group, time `steps` RK4 steps (four stage launches each) of
// every launch candidate -- tile sizes of the small-N kernel, the MFMA, pipelined and
// generic kernels -- and keep the fastest (a candidate replaces the base one only
// when it is at least 3% faster), then re-capture the forward graph.  The state
// the timing runs leave is reset by the next run.
int batch_autotune(Batch* b, int steps) {
    if (!b || !b->impl || steps <= 0) return fail(EH_EINVAL, "batch_autotune: bad arguments");
    BatchImpl* im = b->impl;
    if (im->adj) return fail(EH_EINVAL, "batch_autotune: tune before enabling the adjoint");
    HIPCHK(hipSetDevice(b->ctx->device));
    launch_reset(b);
    im->started = false;
    const StageArgs sa = stage_args(im);
    const PrologueArgs pa = prologue_args(im);
    hipLaunchKernelGGL(prologue, dim3(b->n_inst), dim3(256), 0, im->stream, pa);   // V of step 0
    hipEvent_t e0 = nullptr, e1 = nullptr;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    int rc = EH_OK;
    for (Group& g : im->groups) {
        g.cand_ms.assign(g.cand.size(), 0.0);   // (a single candidate is timed too: the report shows it)
        for (size_t c = 0; c < g.cand.size() && rc == EH_OK; c++) {
            for (int st = 0; st < 4 && rc == EH_OK; st++)   // warm-up step
                rc = launch_cfg(g, g.cand[c], stage_io(sa, im, 0, st), im->stream);
            if (hipEventRecord(e0, im->stream) != hipSuccess) rc = fail(EH_EHIP, "autotune: event record");
            for (int k = 0; k < steps && rc == EH_OK; k++)
                for (int st = 0; st < 4 && rc == EH_OK; st++)
                    rc = launch_cfg(g, g.cand[c], stage_io(sa, im, 0, st), im->stream);
            if (rc == EH_OK && hipEventRecord(e1, im->stream) != hipSuccess) rc = fail(EH_EHIP, "autotune: event record");
            if (rc == EH_OK && hipEventSynchronize(e1) != hipSuccess) rc = fail(EH_EHIP, "autotune: synchronize");
            float ms = 0.0f;
            if (rc == EH_OK && hipEventElapsedTime(&ms, e0, e1) != hipSuccess) rc = fail(EH_EHIP, "autotune: timing");
            g.cand_ms[c] = (double)ms / steps;
        }
        if (rc) break;
        size_t best = 0;
        for (size_t c = 1; c < g.cand.size(); c++)
            if (g.cand_ms[c] < 0.97 * g.cand_ms[best]) best = c;
        g.sel = (int)best;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    return capture_forward(b);
}

// Manual choice of group g's launch candidate c (the tuning report lists both), then re-capture.
int batch_select(Batch* b, int g, int c) {
    if (!b || !b->impl) return fail(EH_EINVAL, "batch_select: null batch");
    BatchImpl* im = b->impl;
    if (im->adj) return fail(EH_EINVAL, "batch_select: select before enabling the adjoint");
    if (g < 0 || g >= (int)im->groups.size()) return fail(EH_ERANGE, "batch_select: no such group");
    if (c < 0 || c >= (int)im->groups[(size_t)g].cand.size()) return fail(EH_ERANGE, "batch_select: no such candidate");
    HIPCHK(hipSetDevice(b->ctx->device));
    HIPCHK(hipStreamSynchronize(im->stream));
    im->groups[(size_t)g].sel = c;
    return capture_forward(b);
}

// One line per group and candidate: "group <i> N<=<n> <kernel> tile=<ge> ms=<t> [*]"
std::string batch_tuning_report(const Batch* b) {
    std::string out;
    if (!b || !b->impl) return out;
    const BatchImpl* im = b->impl;
    for (size_t gi = 0; gi < im->groups.size(); gi++) {
        const Group& g = im->groups[gi];
        for (size_t c = 0; c < g.cand.size(); c++) {
            const LaunchCfg& x = g.cand[c];
            const char* kn = x.nt > 0      ? "stage_small"
                             : x.nt == -32 ? "stage_mfma32"
                             : x.nt == -34 ? "stage_mfma32_tp"
                             : x.nt == -64 ? "stage_mfma64_v1"
                             : x.nt == -65 ? "stage_mfma64"
                             : x.nt == -67 ? "stage_sparse"
                             : x.nt == -69 ? "stage_sparse_rows"
                             : x.nt == -33 ? "stage_mfma_pipe"
                             : x.nt == -128 ? "stage_mfma_wide"
                                            : "stage_generic";
            char line[256];
            std::snprintf(line, sizeof(line),
                          "group %zu cand %zu maxN^2=%d instances=%d %s tile=%d tiles=%d cols=%d stream=%d late=%d ms_per_step=%.6f%s\n",
                          gi, c, g.maxnn, g.inst_count, kn, (x.nt > 0 || x.nt == -128) ? x.ge : 0, x.tile_count,
                          x.cols ? 1 : 0,
                          x.stream ? 1 : 0, x.late ? 1 : 0,
                          g.cand_ms.empty() ? -1.0 : g.cand_ms[c], (int)c == g.sel ? " *" : "");
            out += line;
        }
    }
    return out;
}

int batch_get_rho(Batch* b, int i, double* rho_out) {
    if (!b || i < 0 || i >= b->n_inst || !rho_out) return fail(EH_EINVAL, "batch_get_rho: bad arguments");
    if (!b->impl->ran) return fail(EH_EINVAL, "batch_get_rho: the batch has not run (its state is undefined)");
    const Batch::Info& f = b->info[(size_t)i];
    HIPCHK(hipSetDevice(b->ctx->device));
    HIPCHK(hipStreamSynchronize(b->impl->stream));
    HIPCHK(hipMemcpy(rho_out, b->impl->Y.p + f.state_off, (size_t)f.M * f.N * f.N * sizeof(double2),
                     hipMemcpyDeviceToHost));
    return EH_OK;
}

}  // namespace eh
