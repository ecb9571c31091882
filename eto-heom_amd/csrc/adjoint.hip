// adjoint.hip -- discrete adjoint of the batched RK4 HEOM propagation.
//
// Forward step i (V_i frozen, A_i the hierarchy generator of the RHS):
//   Y1 = y, k1 = A Y1, Y2 = y + h/2 k1, k2 = A Y2, Y3 = y + h/2 k2, k3 = A Y3,
//   Y4 = y + h k3, k4 = A Y4, y' = y + h/6 k1 + h/3 k2 + h/3 k3 + h/6 k4,
// with readouts P_i = Tr(rho_0(y_i) X3) before the step.  For a real loss L and
// the gradient convention  g = dL/dRe z + i dL/dIm z,  the exact reverse step is
//   kb4 = h/6 l'           U4 = A^H kb4     (stage A: input l')
//   kb3 = h/3 l' + h U4    U3 = A^H kb3     (stage B)
//   kb2 = h/3 l' + h/2 U3  U2 = A^H kb2     (stage C)
//   kb1 = h/6 l' + h/2 U2  U1 = A^H kb1     (stage D)
//   l = l' + U1 + U2 + U3 + U4,  then  l[ADO 0][j][k] += conj(X3[k][j]) g_i.
// A^H is the transposed hierarchy: i (V^H l - l V^H), conj(damping), and the
// coupling of ADO m gathers l at m - e_p with the conjugated up-weights of that
// ADO and l at m + e_p with the conjugated down-weights (so the up and down
// tables swap roles; see build of the adjoint tables in adj_enable).
//
// Parameter gradients accumulate sum_i sum_s Re <kb_s, (dA/dtheta) Y_s> on the
// device, per instance:
//   Gt  = sum (Y_n kb_n^H - kb_n^H Y_n)             -> dL/dV = (-i Gt)^H
//   Gq  = sum e_q(t_i) (...)                        -> dL/dX'_q
//   Sg_p = sum_n n_p <kb_n, Y_n>                    -> dL/dgamma_p = -conj(Sg_p)
//   U_p = sum x R_up, A_p = sum y R_dnrow, B_p = sum y R_dncol  -> dL/dalpha, dL/dalpha~
// and the readout term  dL/dX3[k][j] += conj(rho_0[j][k]) g_i.
//
// Memory: the forward run stores y every C steps; the backward replays each
// segment forward from its checkpoint, keeping every stage input of the
// segment, then walks the segment backwards.  Both are graph-captured.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <vector>

#include "engine_impl.hpp"

namespace eh {

struct AdjArgs {
    StageArgs S;             // adjoint coupling tables, conj(alpha), conj(damp), V = V^H, gsign = -1
    const double2* Yf;       // forward stage input paired with this adjoint stage
    const int* fnb_up;       // forward tables (for the coupling-parameter reductions)
    const int* fnb_dn;
    const double* fcx;
    const double* fcy;
    const uint8_t* ado;      // multi-indices in (ADO, p) layout
    const double2* field;    // per step e_q
    double2* gacc;           // per instance accumulators, ncopy copies each
    const long long* gacc_off;
    int ncopy;               // tile t adds into copy t % ncopy (spreads the atomics)
    int ntiles;              // adj_mfma64<true>: the launch's tiles (two per workgroup)
    int half_lds;            // adj_mfma64<true>: bytes of LDS per half workgroup
    long long traj;          // trajectory mode: Yf moves by step * traj elements
    double2* slots;          // small-N tiles: one accumulator block per tile (non-returning atomic adds)
    const long long* slot_off;
    // readout terms, folded into stage 3 of the workgroup holding ADO 0
    const int* pidx;         // per step readout slot (-1: none)
    const double2* R;        // readout operators (X3)
    const double2* gP;       // dL/dP per readout
};


// layout of one instance's accumulator block (complex entries)
__host__ __device__ inline long long gacc_size(int NN, int d) { return 5LL * NN + 4LL * d; }
// [0, NN) Gt | [NN, 4NN) G_q weighted by e_q | [4NN, 5NN) GX3 | Sg[d] | U[d] | A[d] | B[d]
// Each instance holds ncopy such blocks; the host sums them after the backward pass.

__device__ __forceinline__ void atomic_add2(double2* p, double2 v) {
    atomicAdd(&p->x, v.x);
    atomicAdd(&p->y, v.y);
}

__device__ __forceinline__ double2 cconj_mul(double2 a, double2 b) {   // conj(a) * b
    return make_double2(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}

// Stage 3 of step i ends with lambda_i, to which the readout P_i = Tr(rho_0 X3) before
// step i adds dL/dlambda_0[j][k] = conj(X3[k][j]) g_i, g_i = dL/dP_i (and dL/dX3[k][j]
// gets conj(rho_0[j][k]) g_i from the forward state of step i, this stage's Yf).  Only
// 2DES instances (one readout operator) read out P.
__device__ __forceinline__ bool readout_g(const AdjArgs& J, const InstDev& I, int step, double2& g) {
    const int slot = I.nread == 1 ? J.pidx[I.step_off + step] : -1;
    if (slot < 0) return false;
    g = J.gP[I.p_off + slot];
    return true;
}

// the readout term of ADO 0's element (j, k) into the stage-3 update u, and its dL/dX3 term
__device__ __forceinline__ void readout_fold(const AdjArgs& J, const InstDev& I, int ii, int j, int k, double2 rho,
                                             double2 g, double& ur, double& ui) {
    const int N = I.N;
    const double2 t = cconj_mul(J.R[I.r_off + k * N + j], g);
    ur += t.x;
    ui += t.y;
    atomic_add2(J.gacc + J.gacc_off[ii] + 4LL * N * N + k * N + j, cconj_mul(rho, g));
}

// adjoint RK stage update of one element (coefficients of the reverse step above), from
// the stage's input x, lam and ACC (each read where used; no stage uses both lam and ACC,
// and stage 0's lam is its input x).  Two stages write no
// accumulator: stage 0's output is x1 = h/3 lam + h^2/6 u0, so stage 1 rebuilds
// lam + h/6 u0 = 2/3 lam + x1/h from its own input; stage 2's output is
// x3 = h/2 u2 + h/6 lam, so stage 3 adds u2 = 2/h x3 - lam/3 from its own input.  ACC
// holds the classic accumulator less lam/3, so stage 3 reads ACC but no lam (5 state
// writes per step instead of 6).
__device__ __forceinline__ void rk_store_adj(const StageArgs& A, const InstDev& I, size_t gi, double2 x, double2 lam,
                                             double2 acc, double ur, double ui) {
    const int st = A.stage;
    const double h = I.h;
    const double ih = 1.0 / h;
    if (st == 0) {
        const double b = h / 3.0, c = h * h / 6.0;
        A.out[gi] = make_double2(fma(c, ur, b * lam.x), fma(c, ui, b * lam.y));
    } else if (st < 3) {
        const double b = st == 1 ? h / 3.0 : h / 6.0, c = h / 2.0;
        if (st == 1) {
            const double t = 1.0 / 3.0;
            A.acc[gi] = make_double2(fma(t, lam.x, fma(x.x, ih, ur)), fma(t, lam.y, fma(x.y, ih, ui)));
        }
        A.out[gi] = make_double2(fma(c, ur, b * lam.x), fma(c, ui, b * lam.y));
    } else {
        const double t = 2.0 * ih;
        A.out[gi] = make_double2(fma(t, x.x, acc.x + ur), fma(t, x.y, acc.y + ui));
    }
}

// Reductions of the coupling-parameter terms for the ADOs [n0, n0+na) of one
// instance: entries (n, j, rc, pos) over row / column s = j + so, pairing kb_n
// (sKb, scaled by fs) with the forward stage input of the neighbours.
__device__ void reduce_bath(const AdjArgs& J, const InstDev& I, int n0, int na, const double2* sKb, int kstride,
                            double fs, double* sU, int tid, int nthr) {
    const int N = I.N, NN = N * N, K = I.K, Km = I.Km, d = I.d;
    const int per = 2 * K * N;
    for (int e = tid; e < per * na; e += nthr) {
        const int g = e / per, eg = e - g * per;
        const int j = eg / (2 * N), rem = eg - j * 2 * N;
        const int rc = rem / N, pos = rem - rc * N;
        const int s = j + I.so;
        const int ab = rc == 0 ? s * N + pos : pos * N + s;
        double2 kb = sKb[g * kstride + ab];
        kb.x *= fs;
        kb.y *= fs;
        const long long row0 = I.lk_off + (long long)(n0 + g) * d;
        for (int k = 0; k < Km; ++k) {
            const int p = j + k * K;
            const long long q = row0 + p;
            const int mu = J.fnb_up[q], md = J.fnb_dn[q];
            if (mu >= 0) {
                const double2 yu = J.Yf[I.state_off + (long long)mu * NN + ab];
                double2 t = cconj_mul(kb, yu);
                const double x = J.fcx[q] * (rc == 0 ? 1.0 : -1.0);
                // x * (+-i) * conj(kb) * yu
                atomicAdd(&sU[2 * p], -x * t.y);
                atomicAdd(&sU[2 * p + 1], x * t.x);
            }
            if (md >= 0) {
                const double2 yd = J.Yf[I.state_off + (long long)md * NN + ab];
                double2 t = cconj_mul(kb, yd);
                const double y = J.fcy[q] * (rc == 0 ? 1.0 : -1.0);
                double* dst = sU + 2 * d * (rc == 0 ? 1 : 2);   // A_p (row) or B_p (column)
                atomicAdd(&dst[2 * p], -y * t.y);
                atomicAdd(&dst[2 * p + 1], y * t.x);
            }
        }
    }
}

// reduce_bath for one ADO whose kb sits in LDS at row stride LD (adj_mfma64)
__device__ void reduce_bath_ld(const AdjArgs& J, const InstDev& I, int n0, const double2* sKb, int LD, double fs,
                               double* sU, int tid, int nthr) {
    const int N = I.N, NN = N * N, K = I.K, Km = I.Km, d = I.d;
    const int per = 2 * K * N;
    for (int e = tid; e < per; e += nthr) {
        const int j = e / (2 * N), rem = e - j * 2 * N;
        const int rc = rem / N, pos = rem - rc * N;
        const int s = j + I.so;
        const int ab = rc == 0 ? s * N + pos : pos * N + s;
        double2 kb = sKb[rc == 0 ? s * LD + pos : pos * LD + s];
        kb.x *= fs;
        kb.y *= fs;
        const long long row0 = I.lk_off + (long long)n0 * d;
        for (int k = 0; k < Km; ++k) {
            const int p = j + k * K;
            const long long q = row0 + p;
            const int mu = J.fnb_up[q], md = J.fnb_dn[q];
            if (mu >= 0) {
                const double2 yu = J.Yf[I.state_off + (long long)mu * NN + ab];
                const double2 t = cconj_mul(kb, yu);
                const double x = J.fcx[q] * (rc == 0 ? 1.0 : -1.0);
                atomicAdd(&sU[2 * p], -x * t.y);
                atomicAdd(&sU[2 * p + 1], x * t.x);
            }
            if (md >= 0) {
                const double2 yd = J.Yf[I.state_off + (long long)md * NN + ab];
                const double2 t = cconj_mul(kb, yd);
                const double y = J.fcy[q] * (rc == 0 ? 1.0 : -1.0);
                double* dst = sU + 2 * d * (rc == 0 ? 1 : 2);
                atomicAdd(&dst[2 * p], -y * t.y);
                atomicAdd(&dst[2 * p + 1], y * t.x);
            }
        }
    }
}

// The coupling-parameter reductions with one thread per (ADO, bath j, row/column,
// mode k): it sums conj(kb) rho over the N positions of row / column s = j + so
// of the tile's ADO (kb from LDS) against the forward stage input of the up and
// down neighbours, then adds its two results into the LDS sums (x: U_p;  y: A_p
// for rows, B_p for columns; p = j + k K).
// Forward-table values of the thread's first (ADO, bath, row/column, mode) slot,
// loaded at kernel entry (the reduction itself runs after the stage update).
struct BathSlot {
    int mu, md;
    double x, y;
};

__device__ __forceinline__ int bath_slot_q(const InstDev& I, int n0, int t, int& p) {
    const int K = I.K, Km = I.Km;
    const int k = t % Km, r1 = t / Km;
    const int r2 = r1 / 2;
    const int j = r2 % K, g = r2 / K;
    p = j + k * K;
    return (n0 + g) * I.d + p;
}

__device__ __forceinline__ BathSlot bath_slot_load(const AdjArgs& J, const InstDev& I, int n0, int na, int tid) {
    const int nt = na * I.K * 2 * I.Km;
    int p;
    const int t = tid < nt ? tid : 0;
    const long long q = I.lk_off + bath_slot_q(I, n0, t, p);
    BathSlot b;
    b.mu = J.fnb_up[q];
    b.md = J.fnb_dn[q];
    b.x = J.fcx[q];
    b.y = J.fcy[q];
    return b;
}

template <int NT>
__device__ __forceinline__ void reduce_bath_rows(const AdjArgs& J, const InstDev& I, int n0, int na,
                                                 const double2* sKb, double fs, double* sU, int tid, int nthr,
                                                 const BathSlot& first) {
    constexpr int NN = NT * NT;
    const int K = I.K, Km = I.Km, d = I.d;
    const int nt = na * K * 2 * Km;
    for (int t = tid; t < nt; t += nthr) {
        const int k = t % Km, r1 = t / Km;
        const int rc = r1 % 2, r2 = r1 / 2;
        const int j = r2 % K, g = r2 / K;
        const int s = j + I.so, p = j + k * K;
        const long long q = I.lk_off + (long long)(n0 + g) * d + p;
        const bool pre = t == tid;   // the first slot's table values came with the tile's loads
        const int mu = pre ? first.mu : J.fnb_up[q], md = pre ? first.md : J.fnb_dn[q];
        const double x = pre ? first.x : J.fcx[q], y = pre ? first.y : J.fcy[q];
        const double2* kb = sKb + g * NN;
        const double2* Yu = J.Yf + I.state_off + (long long)(mu > 0 ? mu : 0) * NN;
        const double2* Yd = J.Yf + I.state_off + (long long)(md > 0 ? md : 0) * NN;
        // in chunks of up to 8 positions: all loads of a chunk in flight together
        constexpr int CH = NT < 8 ? NT : 8;
        double2 tu = make_double2(0.0, 0.0), td = make_double2(0.0, 0.0);
#pragma unroll 1
        for (int p0 = 0; p0 < NT; p0 += CH) {
            double2 yu[CH], yd[CH];
#pragma unroll
            for (int c = 0; c < CH; ++c) {
                const int pos = p0 + c < NT ? p0 + c : NT - 1;
                const int ab = rc == 0 ? s * NT + pos : pos * NT + s;
                yu[c] = Yu[ab];
                yd[c] = Yd[ab];
            }
#pragma unroll
            for (int c = 0; c < CH; ++c) {
                if (p0 + c >= NT) break;
                const int pos = p0 + c;
                const double2 kv = kb[rc == 0 ? s * NT + pos : pos * NT + s];
                const double2 a = cconj_mul(kv, yu[c]), e = cconj_mul(kv, yd[c]);
                tu.x += a.x;
                tu.y += a.y;
                td.x += e.x;
                td.y += e.y;
            }
        }
        const double sgn = (rc == 0 ? 1.0 : -1.0) * fs;
        if (mu >= 0) {   // x (+-i) conj(kb) yu
            atomicAdd(&sU[2 * p], -sgn * x * tu.y);
            atomicAdd(&sU[2 * p + 1], sgn * x * tu.x);
        }
        if (md >= 0) {
            double* dst = sU + 2 * d * (rc == 0 ? 1 : 2);   // A_p (row) or B_p (column)
            atomicAdd(&dst[2 * p], -sgn * y * td.y);
            atomicAdd(&dst[2 * p + 1], sgn * y * td.x);
        }
    }
}

template <int NT>
__global__ __launch_bounds__(256) void adj_small(AdjArgs J) {
    const StageArgs& A = J.S;
    constexpr int NN = NT * NT;
    constexpr int G = NN >= 512 ? 1 : 512 / NN;
    constexpr int EPT = (G * NN + 255) / 256;
    constexpr int DMAX = 64;
    __shared__ double2 sR[G * NN];
    __shared__ double2 sY[G * NN];
    __shared__ double2 sV[NN];
    extern __shared__ double2 sC[];   // G * 2 K N coupling rows/columns (launch-time size)
    __shared__ double2 sG[G * NN];
    __shared__ double sU[3 * 2 * DMAX];
    __shared__ double sSg[2 * DMAX];
    const int tile = blockIdx.x;
    const int ii = A.tile_inst[tile];
    // a register copy: read once before any store (a reference would be re-read after them)
    const InstDev I = A.inst[ii];
    const int step = *A.step_ctr + A.step_off;
    if (step < 0 || step >= I.nsteps) return;
    J.Yf += (long long)step * J.traj;
    const int a0 = A.tile_ado0[tile], na = A.tile_nado[tile];
    const long long base = I.state_off + (long long)a0 * NN;
    const int nel = na * NN;
    const int st = A.stage;
    // loads first, all unconditional (clamped address, selected result): the
    // neighbour tables, then the tile's inputs, then the neighbour gathers
    const int total = 2 * I.K * NT * na;
    const bool has_cpl = total > 0 && I.Km > 0;
    NbIdx<2, 2> nbx;
    if (has_cpl) gather_idx<2, 2, NT>(A, I, a0, na, threadIdx.x, 256, nbx);
    const bool vreal = A.vreal[I.step_off + step] != 0;
    // tail loads issued with the tile's: forward-table slot of the bath reduction and
    // the occupation n_p of the damping-gradient slot (t = n d + p)
    BathSlot bslot;
    if (has_cpl) bslot = bath_slot_load(J, I, a0, na, threadIdx.x);
    const int dq = na * I.d;
    const int np_pre = dq > 0 ? J.ado[I.lk_off + (long long)a0 * I.d + (threadIdx.x < dq ? threadIdx.x : 0)] : 0;
    // the one state operand of the update besides the input (no stage reads both): lambda
    // in stages 1-2, ACC in stage 3, the input (lambda itself) in stage 0
    const double2* wp = st == 3 ? A.acc : ((st == 1 || st == 2) ? A.y : A.rin);
    double2 rv[EPT], wv[EPT], yf[EPT], gv[EPT];
#pragma unroll
    for (int q = 0; q < EPT; ++q) {
        const int e = threadIdx.x + q * 256;
        const int ec = e < nel ? e : nel - 1;
        rv[q] = A.rin[base + ec];
        wv[q] = wp[base + ec];
        yf[q] = J.Yf[base + ec];
        gv[q] = A.damp[I.ado_off + a0 + ec / NN];
    }
    constexpr int VPT = (NN + 255) / 256;
    double2 vv[VPT];
#pragma unroll
    for (int q = 0; q < VPT; ++q) {
        const int e = threadIdx.x + q * 256;
        vv[q] = A.V[I.v_off + (e < NN ? e : 0)];
    }
    if (has_cpl) {
        gather_data<2, 2, NT>(A, I, na, threadIdx.x, 256, nbx, sC);
        if (total > 2 * 256 || I.Km > 2) gather_rest<2, 2>(A, I, a0, na, threadIdx.x, 256, sC);
    }
    for (int e = threadIdx.x; e < 6 * I.d; e += 256) sU[e] = 0.0;
    for (int e = threadIdx.x; e < 2 * I.d; e += 256) sSg[e] = 0.0;
#pragma unroll
    for (int q = 0; q < VPT; ++q) {
        const int e = threadIdx.x + q * 256;
        if (e < NN) sV[e] = vv[q];
    }
#pragma unroll
    for (int q = 0; q < EPT; ++q) {
        const int e = threadIdx.x + q * 256;
        if (e < nel) {
            sR[e] = rv[q];
            sY[e] = yf[q];
        }
    }
    __syncthreads();
    const double fs = st == 0 ? I.h / 6.0 : 1.0;   // kb = fs * stage input
    const int cpn = 2 * I.K * NT;
    double2 rg = make_double2(0.0, 0.0);
    const bool rdo = st == 3 && a0 == 0 && readout_g(J, I, step, rg);
#pragma unroll
    for (int q = 0; q < EPT; ++q) {
        const int e = threadIdx.x + q * 256;
        if (e >= nel) break;
        const int nl = e / NN, ab = e - nl * NN, ra = ab / NT, cb = ab - ra * NT;
        const double2* R = sR + nl * NN;
        const double2* Y = sY + nl * NN;
        double cr = 0.0, ci = 0.0;   // (V^H l - l V^H)[ra][cb]
        double2 gt = make_double2(0.0, 0.0);
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
        // Gt[ra][cb] = sum_c Y[ra][c] conj(kb[cb][c]) - conj(kb[c][ra]) Y[c][cb]
#pragma unroll
        for (int c = 0; c < NT; ++c) {
            const double2 y1 = Y[ra * NT + c], k1 = R[cb * NT + c];
            const double2 k2 = R[c * NT + ra], y2 = Y[c * NT + cb];
            gt.x += y1.x * k1.x + y1.y * k1.y - (k2.x * y2.x + k2.y * y2.y);
            gt.y += y1.y * k1.x - y1.x * k1.y - (k2.x * y2.y - k2.y * y2.x);
        }
        sG[e] = make_double2(fs * gt.x, fs * gt.y);
        double ur = -ci, ui = cr;   // +i * C'
        const double2 g = gv[q], r = R[ab];
        ur = fma(g.x, r.x, fma(-g.y, r.y, ur));
        ui = fma(g.x, r.y, fma(g.y, r.x, ui));
        const double2 cpl = coupling_of<NT>(I, sC + nl * cpn, ra, cb);
        if (rdo && nl == 0) readout_fold(J, I, ii, ra, cb, Y[ab], rg, ur, ui);
        rk_store_adj(A, I, (size_t)(base + e), r, wv[q], wv[q], ur + cpl.x, ui + cpl.y);
    }
    if (has_cpl) reduce_bath_rows<NT>(J, I, a0, na, sR, fs, sU, threadIdx.x, 256, bslot);
    __syncthreads();
    // this tile's own accumulator block: no other workgroup touches it in this launch, so
    // each entry gets one add per launch in launch order; non-returning atomic adds
    // (measured 1.2 % faster backward than a read-modify-write, whose read the
    // workgroup waits for)
    double2* acc = J.slots + J.slot_off[tile];
    const double2* ef = J.field + 3 * (size_t)(I.step_off + step);
    const double2 e0 = ef[0], e1 = ef[1], e2 = ef[2];
    const bool field_on = e0.x != 0.0 || e0.y != 0.0 || e1.x != 0.0 || e1.y != 0.0 || e2.x != 0.0 || e2.y != 0.0;
    auto add2 = [](double2* p, double2 v) { atomic_add2(p, v); };
    for (int ab = threadIdx.x; ab < NN; ab += 256) {
        double2 gs = make_double2(0.0, 0.0);
        for (int n = 0; n < na; ++n) {
            gs.x += sG[n * NN + ab].x;
            gs.y += sG[n * NN + ab].y;
        }
        add2(acc + ab, gs);
        if (field_on) {   // G_q = sum e_q Gt: nothing to add outside the pulse windows
            add2(acc + NN + ab, cmad(make_double2(0.0, 0.0), e0, gs));
            add2(acc + 2 * NN + ab, cmad(make_double2(0.0, 0.0), e1, gs));
            add2(acc + 3 * NN + ab, cmad(make_double2(0.0, 0.0), e2, gs));
        }
    }
    const int d = I.d;
    // q_n = fs sum_ab conj(kb_n[ab]) Y_n[ab] in two levels (8 partial sums per ADO, from the
    // stage input and the forward state still in LDS), then n_p q_n per (n, p)
    constexpr int QP = NN < 8 ? NN : 8, QC = (NN + QP - 1) / QP;   // na * QP <= G * NN: fits sG
    double2* sPart = sG;   // the Gt partials are consumed above; reuse their LDS
    __syncthreads();
    for (int t = threadIdx.x; t < na * QP; t += 256) {
        const int n = t / QP, part = t - n * QP;
        double2 qs = make_double2(0.0, 0.0);
#pragma unroll
        for (int c = 0; c < QC; ++c) {
            const int ab = part * QC + c;
            if (ab < NN) {
                const double2 qv = cconj_mul(sR[n * NN + ab], sY[n * NN + ab]);
                qs.x += qv.x;
                qs.y += qv.y;
            }
        }
        sPart[t] = make_double2(fs * qs.x, fs * qs.y);
    }
    __syncthreads();
    for (int t = threadIdx.x; t < na * d; t += 256) {   // n_p q_n summed over the tile's ADOs in LDS
        const int n = t / d, p = t - n * d;
        const int np_ = t == (int)threadIdx.x ? np_pre : J.ado[I.lk_off + (long long)(a0 + n) * d + p];
        if (np_ == 0) continue;
        double2 qn = make_double2(0.0, 0.0);
#pragma unroll
        for (int part = 0; part < QP; ++part) {
            qn.x += sPart[n * QP + part].x;
            qn.y += sPart[n * QP + part].y;
        }
        atomicAdd(&sSg[2 * p], np_ * qn.x);
        atomicAdd(&sSg[2 * p + 1], np_ * qn.y);
    }
    __syncthreads();
    for (int t = threadIdx.x; t < 4 * d; t += 256) {   // S_gamma, U, A, B
        const double* src = t < d ? sSg + 2 * t : sU + 2 * (t - d);
        add2(acc + 5 * NN + t, make_double2(src[0], src[1]));
    }
}

// N > 16 outside the MFMA kernel's range: one ADO per workgroup (VALU).  Dynamic LDS: the
// lambda input and Y (G false), the coupling rows/columns, the bath sums.  G: lambda and Y
// are read where they lie (neither is written by the stage: its outputs go to other
// buffers), for N whose 2 N^2 complex do not fit the LDS.
template <bool G>
__global__ __launch_bounds__(256) void adj_generic(AdjArgs J) {
    const StageArgs& A = J.S;
    extern __shared__ double2 smem[];
    const int tile = blockIdx.x;
    const int ii = A.tile_inst[tile];
    const InstDev I = A.inst[ii];   // register copy (read before any store)
    const int step = *A.step_ctr + A.step_off;
    if (step < 0 || step >= I.nsteps) return;
    J.Yf += (long long)step * J.traj;
    const int N = I.N, NN = N * N, d = I.d;
    __shared__ double2 red[256];
    const int a0 = A.tile_ado0[tile];
    const int st = A.stage;
    const long long base = I.state_off + (long long)a0 * NN;
    const double2* sR = G ? A.rin + base : smem;
    const double2* sY = G ? J.Yf + base : smem + NN;
    double2* sC = G ? smem : smem + 2 * NN;
    double* sU = reinterpret_cast<double*>(sC + 2 * I.K * N);
    if (!G) {
        for (int e = threadIdx.x; e < NN; e += blockDim.x) {
            smem[e] = A.rin[base + e];
            smem[NN + e] = J.Yf[base + e];
        }
    }
    for (int e = threadIdx.x; e < 6 * d; e += blockDim.x) sU[e] = 0.0;
    gather_couplings(A, I, a0, 1, sC, threadIdx.x, blockDim.x);
    __syncthreads();
    const double2* V = A.V + I.v_off;
    const bool vreal = A.vreal[I.step_off + step] != 0;
    const double fs = st == 0 ? I.h / 6.0 : 1.0;
    const double2 g = A.damp[I.ado_off + a0];
    double2* acc = J.gacc + J.gacc_off[ii] + (long long)(tile % J.ncopy) * gacc_size(NN, d);
    const double2* ef = J.field + 3 * (size_t)(I.step_off + step);
    const double2 e0 = ef[0], e1 = ef[1], e2 = ef[2];
    const bool field_on = e0.x != 0.0 || e0.y != 0.0 || e1.x != 0.0 || e1.y != 0.0 || e2.x != 0.0 || e2.y != 0.0;
    double2 qsum = make_double2(0.0, 0.0);
    double2 rg = make_double2(0.0, 0.0);
    const bool rdo = st == 3 && a0 == 0 && readout_g(J, I, step, rg);
    for (int ab = threadIdx.x; ab < NN; ab += blockDim.x) {
        const int ra = ab / N, cb = ab - ra * N;
        double cr = 0.0, ci = 0.0;
        double2 gt = make_double2(0.0, 0.0);
        for (int c = 0; c < N; ++c) {
            const double2 vac = V[ra * N + c], vcb = V[c * N + cb];
            const double2 rcb = sR[c * N + cb], rac = sR[ra * N + c];
            if (vreal) {
                cr = fma(vac.x, rcb.x, fma(-rac.x, vcb.x, cr));
                ci = fma(vac.x, rcb.y, fma(-rac.y, vcb.x, ci));
            } else {
                cr = fma(vac.x, rcb.x, fma(-vac.y, rcb.y, cr));
                ci = fma(vac.x, rcb.y, fma(vac.y, rcb.x, ci));
                cr = fma(-rac.x, vcb.x, fma(rac.y, vcb.y, cr));
                ci = fma(-rac.x, vcb.y, fma(-rac.y, vcb.x, ci));
            }
            const double2 y1 = sY[ra * N + c], k1 = sR[cb * N + c];
            const double2 k2 = sR[c * N + ra], y2 = sY[c * N + cb];
            gt.x += y1.x * k1.x + y1.y * k1.y - (k2.x * y2.x + k2.y * y2.y);
            gt.y += y1.y * k1.x - y1.x * k1.y - (k2.x * y2.y - k2.y * y2.x);
        }
        gt = make_double2(fs * gt.x, fs * gt.y);
        atomic_add2(acc + ab, gt);
        if (field_on) {
            atomic_add2(acc + NN + ab, cmad(make_double2(0.0, 0.0), e0, gt));
            atomic_add2(acc + 2 * NN + ab, cmad(make_double2(0.0, 0.0), e1, gt));
            atomic_add2(acc + 3 * NN + ab, cmad(make_double2(0.0, 0.0), e2, gt));
        }
        const double2 qv = cconj_mul(sR[ab], sY[ab]);
        qsum.x += fs * qv.x;
        qsum.y += fs * qv.y;
        double ur = -ci, ui = cr;
        const double2 r = sR[ab];
        ur = fma(g.x, r.x, fma(-g.y, r.y, ur));
        ui = fma(g.x, r.y, fma(g.y, r.x, ui));
        const double2 cpl = coupling_of(I, sC, ra, cb);
        if (rdo) readout_fold(J, I, ii, ra, cb, sY[ab], rg, ur, ui);
        const double2 w = rk_operand(A, (size_t)(base + ab), r);   // lambda (stages 0-2) or ACC
        rk_store_adj(A, I, (size_t)(base + ab), r, w, w, ur + cpl.x, ui + cpl.y);
    }
    red[threadIdx.x] = qsum;
    reduce_bath(J, I, a0, 1, sR, NN, fs, sU, threadIdx.x, blockDim.x);
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            red[threadIdx.x].x += red[threadIdx.x + w].x;
            red[threadIdx.x].y += red[threadIdx.x + w].y;
        }
        __syncthreads();
    }
    const double2 qn = red[0];
    for (int p = threadIdx.x; p < d; p += blockDim.x) {
        const int np_ = J.ado[I.lk_off + (long long)a0 * d + p];
        if (np_) atomic_add2(acc + 5 * NN + p, make_double2(np_ * qn.x, np_ * qn.y));
    }
    for (int t = threadIdx.x; t < 3 * d; t += blockDim.x) {
        const int w = t / d, p = t - w * d;
        atomic_add2(acc + 5 * NN + (1 + w) * d + p, make_double2(sU[2 * (w * d + p)], sU[2 * (w * d + p) + 1]));
    }
}

// 32 < N <= 64: one ADO per 512 threads (8 waves, each two 16 x 16 tiles of one tile row, as
// stage_mfma64).  LDS holds the adjoint stage input l (N x (N+1) complex), the coupling rows /
// columns and the bath sums; the forward stage input Y is read through the caches.  On FP64
// MFMA:
//   C' = V^H l - l V^H                      (the stage update u = i C' + conj(damp) l + couplings)
//   Gt = Y K^H - K^H Y,  K = fs l           (dL/dV, added into the instance's accumulator copy)
// with the element-wise terms, the readout fold, q_n and the bath reductions as adj_generic.
// PAIR: a 1024-thread workgroup runs two ADOs, one per half (tiles p and p + 8 of the
// XCD-ordered list: one XCD's share, consecutive ADOs of an instance).  When both belong to one
// instance, the second half hands its Gt to the first through its own (then unused) LDS and
// the first adds both with one set of atomics, which cost by their number.  Every barrier is
// reached by both halves; a half without a tile skips the work between them.
template <bool PAIR>
__global__ __launch_bounds__(PAIR ? 1024 : 512, PAIR ? 1 : 2) void adj_mfma64(AdjArgs J) {
    const StageArgs& A = J.S;
    extern __shared__ double2 smem[];
    // the half: uniform in each wave, so a scalar (the instance data then stays in SGPRs)
    const int half = PAIR ? __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 9)) : 0;
    const int tid = (int)(threadIdx.x & 511);
    const int tile0 = PAIR ? (int)((blockIdx.x >> 3) * 16 + (blockIdx.x & 7)) : (int)blockIdx.x;
    const int tile = tile0 + 8 * half;
    const int step = *A.step_ctr + A.step_off;
    // both halves' instances (clamped reads for a missing second tile)
    const int T = PAIR ? J.ntiles : (int)gridDim.x;
    const int iiA = A.tile_inst[min(tile0, T - 1)], iiB = A.tile_inst[min(tile0 + 8, T - 1)];
    const int ii = half ? iiB : iiA;
    const InstDev& I = A.inst[ii];
    const bool valid = tile < T && step >= 0 && step < I.nsteps;   // uniform in the half
    const bool same = PAIR && tile0 + 8 < T && iiA == iiB && step >= 0 && step < I.nsteps;
    if (!PAIR && !valid) return;
    J.Yf += (long long)step * J.traj;
    const int N = I.N, NN = N * N, d = I.d, LD = N + 1;
    double2* const sbase = reinterpret_cast<double2*>(reinterpret_cast<char*>(smem) + (size_t)half * J.half_lds);
    double2* sR = sbase;                                              // l, N x (N+1)
    double2* sC = sbase + N * LD;                                     // 2 K N coupling rows/columns
    double* sU = reinterpret_cast<double*>(sC + 2 * I.K * N);         // U, A, B (6 d doubles)
    __shared__ double2 red[2][8];
    const int a0 = A.tile_ado0[min(tile, T - 1)];
    const int st = A.stage;
    const long long base = I.state_off + (long long)a0 * NN;
    const double2* Yf = J.Yf + base;
    if (valid) {
        for (int e = tid; e < NN; e += 512) {
            const int r = e / N, c = e - r * N;
            sR[r * LD + c] = A.rin[base + e];
        }
        for (int e = tid; e < 6 * d; e += 512) sU[e] = 0.0;
        gather_couplings<true>(A, I, a0, 1, sC, tid, 512);
    }
    __syncthreads();
    const double2* V = A.V + I.v_off;   // V^H of this step
    const bool vreal = A.vreal[I.step_off + min(max(step, 0), I.nsteps - 1)] != 0;
    const double fs = st == 0 ? I.h / 6.0 : 1.0;
    const int lane = tid & 63, wave = tid >> 6;
    const int li = lane & 15, lk = lane >> 4;
    const int ti = wave >> 1, tj0 = (wave & 1) * 2;
    const int ar = ti * 16 + li;
    const int kmax = (N + 3) & ~3;
    dbl4 cr[2], ci[2];
#pragma unroll
    for (int t = 0; t < 2; ++t) {
        cr[t] = dbl4{0.0, 0.0, 0.0, 0.0};
        ci[t] = dbl4{0.0, 0.0, 0.0, 0.0};
    }
    if (valid) {
        const double2 g = A.damp[I.ado_off + a0];
        // C' = V^H l - l V^H
        if (vreal) {
#pragma unroll 2
            for (int k0 = 0; k0 < kmax; k0 += 4) {
                const int kk = k0 + lk;
                const double va = gl_at(V, ar, kk, N).x;
                const double2 ra = lds_at(sR, ar, kk, N, LD);
#pragma unroll
                for (int t = 0; t < 2; ++t) {
                    const int bc = (tj0 + t) * 16 + li;
                    const double vb = gl_at(V, kk, bc, N).x;
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
                const double2 va = gl_at(V, ar, kk, N);
                const double2 ra = lds_at(sR, ar, kk, N, LD);
#pragma unroll
                for (int t = 0; t < 2; ++t) {
                    const int bc = (tj0 + t) * 16 + li;
                    const double2 vb = gl_at(V, kk, bc, N);
                    const double2 rb = lds_at(sR, kk, bc, N, LD);
                    cr[t] = mfma4(va.x, rb.x, cr[t]);
                    ci[t] = mfma4(va.x, rb.y, ci[t]);
                    cr[t] = mfma4(-va.y, rb.y, cr[t]);
                    ci[t] = mfma4(va.y, rb.x, ci[t]);
                    cr[t] = mfma4(ra.x, -vb.x, cr[t]);
                    ci[t] = mfma4(ra.y, -vb.x, ci[t]);
                    cr[t] = mfma4(ra.y, vb.y, cr[t]);
                    ci[t] = mfma4(ra.x, -vb.y, ci[t]);
                }
            }
        }
        // the stage update of the lane's 8 elements
        double2 rg = make_double2(0.0, 0.0);
        const bool rdo = st == 3 && a0 == 0 && readout_g(J, I, step, rg);
        const double2* wp = st == 3 ? A.acc : ((st == 1 || st == 2) ? A.y : A.rin);
        double2 wv[2][4];
#pragma unroll
        for (int t = 0; t < 2; ++t) {
            const int col = (tj0 + t) * 16 + li;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = ti * 16 + lk + 4 * r;
                const bool in = row < N && col < N;
                wv[t][r] = wp[(size_t)(base + (in ? row * N + col : 0))];
            }
        }
        double2 qsum = make_double2(0.0, 0.0);
#pragma unroll
        for (int t = 0; t < 2; ++t) {
            const int col = (tj0 + t) * 16 + li;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = ti * 16 + lk + 4 * r;
                if (row < N && col < N) {
                    const double2 x = sR[row * LD + col];
                    double ur = -ci[t][r], ui = cr[t][r];   // +i * C'
                    ur = fma(g.x, x.x, fma(-g.y, x.y, ur));
                    ui = fma(g.x, x.y, fma(g.y, x.x, ui));
                    const double2 cpl = coupling_of(I, sC, row, col);
                    const double2 yv = Yf[row * N + col];
                    const double2 qv = cconj_mul(x, yv);
                    qsum.x += fs * qv.x;
                    qsum.y += fs * qv.y;
                    if (rdo) readout_fold(J, I, ii, row, col, yv, rg, ur, ui);
                    rk_store_adj(A, I, (size_t)(base + row * N + col), x, wv[t][r], wv[t][r], ur + cpl.x,
                                 ui + cpl.y);
                }
            }
        }
        // q_n = fs sum conj(l) Y over the ADO (wave sums now, n_p q_n after the barrier); the
        // bath reductions into sU
        for (int o = 32; o > 0; o >>= 1) {
            qsum.x += __shfl_xor(qsum.x, o, 64);
            qsum.y += __shfl_xor(qsum.y, o, 64);
        }
        if (lane == 0) red[half][wave] = qsum;
        reduce_bath_ld(J, I, a0, sR, LD, fs, sU, tid, 512);
        // Gt = Y K^H - K^H Y (K = fs l): P1[a][b] = sum_c Y[a][c] conj(K[b][c]),
        // P2[a][b] = sum_c conj(K[c][a]) Y[c][b]
#pragma unroll
        for (int t = 0; t < 2; ++t) {
            cr[t] = dbl4{0.0, 0.0, 0.0, 0.0};
            ci[t] = dbl4{0.0, 0.0, 0.0, 0.0};
        }
#pragma unroll 2
        for (int k0 = 0; k0 < kmax; k0 += 4) {
            const int kk = k0 + lk;
            const double2 y1 = gl_at(Yf, ar, kk, N);        // Y[a][c]   (A operand)
            const double2 k2 = lds_at(sR, kk, ar, N, LD);   // l[c][a]   (A operand, conjugated)
#pragma unroll
            for (int t = 0; t < 2; ++t) {
                const int bc = (tj0 + t) * 16 + li;
                const double2 k1 = lds_at(sR, bc, kk, N, LD);   // l[b][c]  (B operand, conjugated)
                const double2 y2 = gl_at(Yf, kk, bc, N);        // Y[c][b]  (B operand)
                cr[t] = mfma4(y1.x, k1.x, cr[t]);
                ci[t] = mfma4(-y1.x, k1.y, ci[t]);
                cr[t] = mfma4(y1.y, k1.y, cr[t]);
                ci[t] = mfma4(y1.y, k1.x, ci[t]);
                cr[t] = mfma4(-k2.x, y2.x, cr[t]);
                ci[t] = mfma4(-k2.x, y2.y, ci[t]);
                cr[t] = mfma4(-k2.y, y2.y, cr[t]);
                ci[t] = mfma4(k2.y, y2.x, ci[t]);
            }
        }
    }
    __syncthreads();   // l, the couplings and sU are read; red and sU are written
    if (PAIR && same) {
        // the second half's Gt into its own LDS, the first adds it to its own
        if (half == 1) {
#pragma unroll
            for (int t = 0; t < 2; ++t) {
                const int col = (tj0 + t) * 16 + li;
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int row = ti * 16 + lk + 4 * r;
                    if (row < N && col < N) sR[row * LD + col] = make_double2(cr[t][r], ci[t][r]);
                }
            }
        }
        __syncthreads();
        if (half == 0) {
            const double2* sRB = reinterpret_cast<const double2*>(reinterpret_cast<const char*>(smem) + J.half_lds);
#pragma unroll
            for (int t = 0; t < 2; ++t) {
                const int col = (tj0 + t) * 16 + li;
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int row = ti * 16 + lk + 4 * r;
                    if (row < N && col < N) {
                        const double2 o = sRB[row * LD + col];
                        cr[t][r] += o.x;
                        ci[t][r] += o.y;
                    }
                }
            }
        }
    }
    if (!valid) return;   // no barrier follows
    double2* acc = J.gacc + J.gacc_off[ii] + (long long)(tile % J.ncopy) * gacc_size(NN, d);
    if (!(PAIR && same && half == 1)) {
        const double2* ef = J.field + 3 * (size_t)(I.step_off + step);
        const double2 e0 = ef[0], e1 = ef[1], e2 = ef[2];
        const bool field_on = e0.x != 0.0 || e0.y != 0.0 || e1.x != 0.0 || e1.y != 0.0 || e2.x != 0.0 || e2.y != 0.0;
#pragma unroll
        for (int t = 0; t < 2; ++t) {
            const int col = (tj0 + t) * 16 + li;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = ti * 16 + lk + 4 * r;
                if (row < N && col < N) {
                    const int ab = row * N + col;
                    const double2 gt = make_double2(fs * cr[t][r], fs * ci[t][r]);
                    atomic_add2(acc + ab, gt);
                    if (field_on) {
                        atomic_add2(acc + NN + ab, cmad(make_double2(0.0, 0.0), e0, gt));
                        atomic_add2(acc + 2 * NN + ab, cmad(make_double2(0.0, 0.0), e1, gt));
                        atomic_add2(acc + 3 * NN + ab, cmad(make_double2(0.0, 0.0), e2, gt));
                    }
                }
            }
        }
    }
    double2 qn = make_double2(0.0, 0.0);
#pragma unroll
    for (int w = 0; w < 8; ++w) {
        qn.x += red[half][w].x;
        qn.y += red[half][w].y;
    }
    for (int p = tid; p < d; p += 512) {
        const int np_ = J.ado[I.lk_off + (long long)a0 * d + p];
        if (np_) atomic_add2(acc + 5 * NN + p, make_double2(np_ * qn.x, np_ * qn.y));
    }
    for (int t = tid; t < 3 * d; t += 512) {
        const int w = t / d, p = t - w * d;
        atomic_add2(acc + 5 * NN + (1 + w) * d + p, make_double2(sU[2 * (w * d + p)], sU[2 * (w * d + p) + 1]));
    }
}

struct AdjVhArgs {
    const InstDev* inst;
    const int* step_ctr;
    const double2* H;      // V of a step from H, X'_q and the field (as the forward builds it)
    const double2* Xp;
    const double2* field;
    double2* VH;           // C copies, copy k for step *step_ctr + k
    long long vstride;     // complex elements per copy
};

// V_i^H = conj(V_i)^T of the C steps of a backward segment, one launch per segment
// (workgroup (instance, k) builds step *step_ctr + k), for the adjoint stages of those
// steps; V is built from H, X'_q and the field exactly as the forward builds it.
__global__ __launch_bounds__(256) void adj_vh(AdjVhArgs P) {
    const int ii = blockIdx.x, k = blockIdx.y;
    const InstDev& I = P.inst[ii];
    const int N = I.N, NN = N * N;
    const int step = *P.step_ctr + k;
    if (step < 0 || step >= I.nsteps) return;
    const double2* e = P.field + 3 * (size_t)(I.step_off + step);
    const double2 e0 = e[0], e1 = e[1], e2 = e[2];
    const double2* H = P.H + I.v_off;
    const double2* X = P.Xp + 3 * (size_t)I.v_off;
    double2* out = P.VH + (long long)k * P.vstride + I.v_off;
    for (int q = threadIdx.x; q < NN; q += blockDim.x) {
        const int a = q / N, b = q - a * N;
        const double2 v = v_elem(H, X, NN, e0, e1, e2, b * N + a);
        out[q] = make_double2(v.x, -v.y);
    }
}

// y -> checkpoint slot (*ctr / C) + 1, at the end of a C-step forward chunk
__global__ void ckpt_store(const double2* y, double2* ckpt, long long total, const int* ctr, int C, int nslots) {
    const int slot = *ctr / C + 1;
    if (slot >= nslots) return;
    double2* dst = ckpt + (long long)slot * total;
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (long long)gridDim.x * blockDim.x)
        dst[e] = y[e];
}

// checkpoint slot (*ctr / C) -> dst, at the start of a backward segment
__global__ void ckpt_load(const double2* ckpt, double2* dst, long long total, const int* ctr, int C) {
    const double2* src = ckpt + (long long)(*ctr / C) * total;
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (long long)gridDim.x * blockDim.x)
        dst[e] = src[e];
}

// trajectory mode: the state each instance ended in (slot 4 nsteps) -> Y
__global__ void traj_final(const double2* traj, double2* y, const InstDev* inst, long long slot_elems) {
    const InstDev& I = inst[blockIdx.x];
    const long long n = (long long)I.M * I.N * I.N;
    const double2* src = traj + 4LL * I.nsteps * slot_elems + I.state_off;
    for (long long e = threadIdx.x; e < n; e += blockDim.x) y[I.state_off + e] = src[e];
}

__global__ void fill_zero(double2* p, long long n) {
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (long long)gridDim.x * blockDim.x)
        p[e] = make_double2(0.0, 0.0);
}

__global__ void set_ctr(int* ctr, int v) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *ctr = v;
}

// ---------------------------------------------------------------- host side

struct AdjState {
    int C = 0;                 // checkpoint interval (steps per segment)
    int ncopy = 32;            // accumulator copies per instance (N > 16 kernels: atomics)
    DevBuf<double2> slots;     // small-N tiles: one accumulator block each
    DevBuf<long long> slot_off;
    std::vector<long long> h_slot_off;   // per tile (-1: not a small-N tile)
    std::vector<double2> h_slots;
    int nseg = 0;
    long long total = 0;       // state elements of the batch
    bool trajm = false;        // trajectory mode: every stage input of the run is kept, no replay
    DevBuf<double2> traj;      // trajectory mode: (4 max_steps + 1) states
    DevBuf<double2> ckpt;      // (nseg + 1) states
    DevBuf<double2> stg;       // (C + 1) x 4 stage inputs of the current segment
    DevBuf<double2> acc_f;     // forward accumulator scratch of the replay
    DevBuf<double2> lam, aacc, q1, q2, VH, gP, gacc;
    DevBuf<int> nbA_up, nbA_dn;
    DevBuf<double> cxA, cyA;
    DevBuf<double2> alA, altA, dampA;
    DevBuf<uint8_t> ado;
    DevBuf<long long> gacc_off;
    std::vector<long long> h_gacc_off;
    hipGraph_t gfwd = nullptr, gbwd = nullptr;
    hipGraphExec_t xfwd = nullptr, xbwd = nullptr;
    bool have_grad = false;
    std::vector<double2> h_gacc;
    ~AdjState() {
        if (xfwd) (void)hipGraphExecDestroy(xfwd);
        if (gfwd) (void)hipGraphDestroy(gfwd);
        if (xbwd) (void)hipGraphExecDestroy(xbwd);
        if (gbwd) (void)hipGraphDestroy(gbwd);
    }
};

void adj_destroy(AdjState* a) { delete a; }

// dynamic LDS of adj_generic for a group: rho and lambda (N^2 each; none with G), the
// coupling rows/columns (2 K N), six d-vectors of bath sums
static size_t adj_generic_lds(const Group& g, bool G = false) {
    return (size_t)((G ? 0 : 2 * g.maxnn) + g.maxkn) * sizeof(double2) + (size_t)6 * g.maxd * sizeof(double);
}

// dynamic LDS of adj_mfma64: l at stride N + 1, the coupling rows/columns, six d-vectors
static size_t adj_mfma64_lds(const Group& g) {
    int mn = 1;
    while (mn * mn < g.maxnn) mn++;
    return (size_t)(mn * (mn + 1) + g.maxkn) * sizeof(double2) + (size_t)6 * g.maxd * sizeof(double);
}

// the adjoint kernel of a group: adj_small for N <= 16, adj_mfma64 where the forward's
// base kernel is stage_mfma64 (32 < N <= 64), adj_generic otherwise
static bool adj_uses_mfma64(const Group& g) { return g.nt == -65; }

static int launch_adj(const Group& g, const AdjArgs& base, hipStream_t s) {
    AdjArgs a = base;
    a.S.tile_inst += g.tile_begin;
    a.S.tile_ado0 += g.tile_begin;
    a.S.tile_nado += g.tile_begin;
    a.slot_off += g.tile_begin;
    const dim3 grid(g.tile_count), block(256);
    switch (g.nt) {
#define EH_CASE(n) \
    case n: hipLaunchKernelGGL(adj_small<n>, grid, block, g.shmem, s, a); break;
        EH_CASE(1) EH_CASE(2) EH_CASE(3) EH_CASE(4) EH_CASE(5) EH_CASE(6) EH_CASE(7) EH_CASE(8)
        EH_CASE(9) EH_CASE(10) EH_CASE(11) EH_CASE(12) EH_CASE(13) EH_CASE(14) EH_CASE(15) EH_CASE(16)
#undef EH_CASE
        default: {
            // every N > 16 group has one ADO per tile
            if (adj_uses_mfma64(g)) {
                const size_t lds = adj_mfma64_lds(g);
                if (g.adj_pair) {   // two ADOs per 1024-thread workgroup (see adj_mfma64)
                    a.ntiles = g.tile_count;
                    a.half_lds = (int)((lds + 15) & ~(size_t)15);
                    hipLaunchKernelGGL(adj_mfma64<true>, dim3((g.tile_count + 15) / 16 * 8), dim3(1024),
                                       2 * (size_t)a.half_lds, s, a);
                } else {
                    hipLaunchKernelGGL(adj_mfma64<false>, grid, dim3(512), lds, s, a);
                }
            }
            else if (g.adj_global)
                hipLaunchKernelGGL(adj_generic<true>, grid, block, adj_generic_lds(g, true), s, a);
            else
                hipLaunchKernelGGL(adj_generic<false>, grid, block, adj_generic_lds(g), s, a);
            break;
        }
    }
    HIPCHK(hipGetLastError());
    return EH_OK;
}

int batch_enable_adjoint(Batch* b, int C) {
    if (!b || !b->impl) return fail(EH_EINVAL, "enable_adjoint: null batch");
    BatchImpl* im = b->impl;
    if (im->adj) return fail(EH_EINVAL, "adjoint already enabled");
    if (C <= 0) return fail(EH_EINVAL, "checkpoint interval must be positive");
    HIPCHK(hipSetDevice(b->ctx->device));
    std::unique_ptr<AdjState> a(new AdjState());
    a->C = std::min(C, std::max(1, b->max_steps));
    if (const char* ev = std::getenv("ETOHEOM_ADJ_COPIES")) a->ncopy = std::max(1, std::atoi(ev));
    a->nseg = (b->max_steps + a->C - 1) / a->C;
    a->total = b->total_elems;
    // adjoint tables: up <- down, down <- up, with the coefficients of the partner link
    const size_t nq = im->hup.size();
    std::vector<int> aup(nq), adn(nq);
    std::vector<double> acx(nq, 0.0), acy(nq, 0.0);
    std::vector<double2> aal(im->hal.size()), aalt(im->halt.size()), adamp(im->hdamp.size());
    std::vector<long long> goff;
    long long gtot = 0;
    std::vector<long long> done;   // instances share hierarchy tables: build each once
    for (int i = 0; i < b->n_inst; i++) {
        const InstDev& I = im->hinst[(size_t)i];
        const bool seen = std::find(done.begin(), done.end(), I.lk_off) != done.end();
        if (!seen) done.push_back(I.lk_off);
        for (int m = 0; m < (seen ? 0 : I.M); m++)
            for (int p = 0; p < I.d; p++) {
                const size_t q = (size_t)(I.lk_off + (long long)m * I.d + p);
                const int j = p % I.K, k = p / I.K;
                const double2 al = im->hal[(size_t)I.al_off + j * I.Km + k];
                const double aabs = std::hypot(al.x, al.y);
                const int mp = im->hado[q];
                aup[q] = im->hdn[q];
                adn[q] = im->hup[q];
                if (aup[q] >= 0) acx[q] = std::sqrt(mp * aabs);          // x of the link (m - e_p) -> m
                if (adn[q] >= 0) acy[q] = std::sqrt((mp + 1) / aabs);    // y of the link (m + e_p) -> m
            }
        goff.push_back(gtot);
        gtot += (long long)a->ncopy * gacc_size(I.N * I.N, I.d);
        if (I.d > 64 && I.N <= 16) return fail(EH_ERANGE, "adjoint: hierarchy dimension above 64");
    }
    for (size_t e = 0; e < aal.size(); e++) {
        aal[e] = make_double2(im->hal[e].x, -im->hal[e].y);
        aalt[e] = make_double2(im->halt[e].x, -im->halt[e].y);
    }
    for (size_t e = 0; e < adamp.size(); e++) adamp[e] = make_double2(im->hdamp[e].x, -im->hdamp[e].y);
    a->h_gacc_off = goff;
    // one accumulator block per tile of the small-N groups (their base tiling)
    a->h_slot_off.assign(im->htile_inst.size(), -1);
    long long stot = 0;
    for (const Group& g : im->groups) {
        if (g.nt < 1 || g.nt > 16) continue;
        for (int t = g.tile_begin; t < g.tile_begin + g.tile_count; t++) {
            const InstDev& I = im->hinst[(size_t)im->htile_inst[(size_t)t]];
            a->h_slot_off[(size_t)t] = stot;
            stot += gacc_size(I.N * I.N, I.d);
        }
    }
    // the N > 16 adjoint kernels stage lambda (adj_generic: also Y) and the coupling
    // rows/columns in LDS; adj_generic reads lambda and Y in place where they do not fit
    for (Group& g : im->groups) {
        if (g.nt >= 1 && g.nt <= 16) continue;
        const bool mf = adj_uses_mfma64(g);
        g.adj_global = !mf && (long long)adj_generic_lds(g) > b->ctx->lds_per_block;
        const size_t sh = mf ? adj_mfma64_lds(g) : adj_generic_lds(g, g.adj_global);
        if ((long long)sh > b->ctx->lds_per_block)
            return fail(EH_ESHAPE, "adjoint: an N > 16 instance needs " + std::to_string(sh) + " bytes of LDS per " +
                                       "workgroup (" + (mf ? "N (N+1) + " : "") + "2 K N complex + 48 d); " +
                                       "the device has " + std::to_string(b->ctx->lds_per_block));
        if (sh > 65536)
            HIPCHK(hipFuncSetAttribute(mf ? (const void*)adj_mfma64<false>
                                          : (g.adj_global ? (const void*)adj_generic<true> : (const void*)adj_generic<false>),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh));
        // two ADOs per workgroup where both halves' LDS fits (ETOHEOM_ADJ_NO_PAIR=1: one)
        const size_t sh2 = 2 * ((sh + 15) & ~(size_t)15);
        const char* np = std::getenv("ETOHEOM_ADJ_NO_PAIR");
        g.adj_pair = mf && (long long)sh2 <= b->ctx->lds_per_block && !(np && np[0] == '1');
        if (g.adj_pair && sh2 > 65536)
            HIPCHK(hipFuncSetAttribute((const void*)adj_mfma64<true>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                       (int)sh2));
    }
    int rc;
    const size_t T = (size_t)a->total;
    // Trajectory mode when every stage input of the run fits in half of the free
    // HBM: the forward writes them into per-step slots (its stage outputs are the
    // next stage's inputs, so this costs no extra traffic) and the backward needs
    // no recomputation.  Otherwise checkpoints every C steps and a replay per segment.
    {
        size_t fr = 0, tot = 0;
        HIPCHK(hipMemGetInfo(&fr, &tot));
        const double need = (double)T * sizeof(double2) * (4.0 * b->max_steps + 1.0);
        const char* ev = std::getenv("ETOHEOM_ADJ_REPLAY");
        a->trajm = !(ev && ev[0] == '1') && need < 0.5 * (double)fr;
    }
    if ((rc = a->nbA_up.upload(aup)) || (rc = a->nbA_dn.upload(adn)) || (rc = a->cxA.upload(acx)) ||
        (rc = a->cyA.upload(acy)) || (rc = a->alA.upload(aal)) || (rc = a->altA.upload(aalt)) ||
        (rc = a->dampA.upload(adamp)) || (rc = a->ado.upload(im->hado)) || (rc = a->gacc_off.upload(goff)) ||
        (a->trajm ? (rc = a->traj.alloc(T * (4 * (size_t)b->max_steps + 1)))
                  : ((rc = a->ckpt.alloc(T * (size_t)(a->nseg + 1))) || (rc = a->stg.alloc(T * 4 * (size_t)(a->C + 1))) ||
                     (rc = a->acc_f.alloc(T)))) ||
        (rc = a->lam.alloc(T)) || (rc = a->aacc.alloc(T)) || (rc = a->q1.alloc(T)) || (rc = a->q2.alloc(T)) ||
        (rc = a->VH.alloc((size_t)a->C * (im->V.n / 2))) || (rc = a->gP.alloc(im->P.n)) || (rc = a->gacc.alloc((size_t)gtot)) ||
        (rc = a->slots.alloc((size_t)stot)) || (rc = a->slot_off.upload(a->h_slot_off)))
        return rc;
    const int blocks = (int)std::min<long long>(std::max<long long>(1, ((long long)T + 255) / 256), 65536);
    hipStream_t s = im->stream;
    StageArgs fw = stage_args(im);
    PrologueArgs pa = prologue_args(im);

    // trajectory slot st of the current step (the kernels add step * 4T)
    auto tslot = [&](int st) { return a->traj.p + (size_t)st * T; };
    if (a->trajm) {
        pa.y = tslot(0);
        pa.traj = 4 * (long long)T;
    }
    // forward: C fused steps (readouts in stage 0, the next step's V in stage 3, as the
    // forward graph does; with the persistent pipelined kernel, which has no fused tail,
    // a prologue launch per step instead) + (checkpoint store) + tick
    bool fused = true;
    for (const Group& g : im->groups) fused = fused && g.cand[(size_t)g.sel].nt != -33;
    HIPCHK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    for (int k = 0; k < a->C; k++) {
        if (!fused) {
            pa.step_off = k;
            hipLaunchKernelGGL(prologue, dim3(b->n_inst), dim3(256), 0, s, pa);
        }
        for (int st = 0; st < 4; st++) {
            StageArgs x = stage_io(fw, im, k, st);
            x.fused = fused ? 1 : 0;
            if (a->trajm) {   // Y_i, S1, S2, S1' of step i in slots 4i .. 4i+3; Y_{i+1} in 4i+4
                x.y = tslot(0);
                x.rin = tslot(st);
                x.out = tslot(st + 1);
                x.traj = 4 * (long long)T;
            }
            for (const Group& g : im->groups)
                if ((rc = launch_stage(g, x, s))) {
                    hipGraph_t junk;
                    (void)hipStreamEndCapture(s, &junk);
                    return rc;
                }
        }
    }
    if (!a->trajm)
        hipLaunchKernelGGL(ckpt_store, dim3(blocks), dim3(256), 0, s, im->Y.p, a->ckpt.p, (long long)T,
                           im->step_ctr.p, a->C, a->nseg + 1);
    hipLaunchKernelGGL(tick, dim3(1), dim3(1), 0, s, im->step_ctr.p, a->C);
    HIPCHK(hipStreamEndCapture(s, &a->gfwd));
    HIPCHK(hipGraphInstantiate(&a->xfwd, a->gfwd, nullptr, nullptr, 0));

    // backward segment: replay from the checkpoint keeping every stage input, then reverse
    auto stg = [&](int k, int st) { return a->stg.p + ((size_t)k * 4 + st) * T; };
    AdjArgs aj;
    std::memset(&aj, 0, sizeof(aj));
    aj.S = fw;
    aj.S.nb_up = a->nbA_up.p;
    aj.S.nb_dn = a->nbA_dn.p;
    aj.S.cx = a->cxA.p;
    aj.S.cy = a->cyA.p;
    aj.S.alpha = a->alA.p;
    aj.S.alpha_t = a->altA.p;
    aj.S.damp = a->dampA.p;
    aj.S.gsign = -1.0;
    aj.S.V = a->VH.p;
    aj.fnb_up = im->nb_up.p;
    aj.fnb_dn = im->nb_dn.p;
    aj.fcx = im->cx.p;
    aj.fcy = im->cy.p;
    aj.ado = a->ado.p;
    aj.field = im->field.p;
    aj.gacc = a->gacc.p;
    aj.gacc_off = a->gacc_off.p;
    aj.ncopy = a->ncopy;
    aj.slots = a->slots.p;
    aj.slot_off = a->slot_off.p;
    aj.pidx = im->pidx.p;
    aj.R = im->R.p;
    aj.gP = a->gP.p;
    AdjVhArgs ar;
    std::memset(&ar, 0, sizeof(ar));
    ar.inst = im->inst.p;
    ar.step_ctr = im->step_ctr.p;
    ar.H = im->H.p;
    ar.Xp = im->Xp.p;
    ar.field = im->field.p;
    ar.VH = a->VH.p;
    ar.vstride = (long long)(im->V.n / 2);
    aj.traj = a->trajm ? 4 * (long long)T : 0;
    pa.readout = 0;   // replays must not overwrite the forward readouts
    HIPCHK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    if (!a->trajm)
        hipLaunchKernelGGL(ckpt_load, dim3(blocks), dim3(256), 0, s, a->ckpt.p, stg(0, 0), (long long)T,
                           im->step_ctr.p, a->C);
    for (int k = 0; k < a->C && !a->trajm; k++) {
        pa.step_off = k;
        hipLaunchKernelGGL(prologue, dim3(b->n_inst), dim3(256), 0, s, pa);
        for (int st = 0; st < 4; st++) {
            StageArgs x = fw;
            x.step_off = k;
            x.stage = st;
            x.y = stg(k, 0);
            x.acc = a->acc_f.p;
            x.rin = stg(k, st);
            x.out = st == 3 ? stg(k + 1, 0) : stg(k, st + 1);
            for (const Group& g : im->groups)
                if ((rc = launch_stage(g, x, s))) {
                    hipGraph_t junk;
                    (void)hipStreamEndCapture(s, &junk);
                    return rc;
                }
        }
    }
    // V^H of every step of the segment (copy k for step k), then the steps in reverse; the
    // readout terms of step k are folded into its adjoint stage 3
    hipLaunchKernelGGL(adj_vh, dim3(b->n_inst, a->C), dim3(256), 0, s, ar);
    for (int k = a->C - 1; k >= 0; k--) {
        for (int st = 0; st < 4; st++) {
            AdjArgs x = aj;
            x.S.V = a->VH.p + (size_t)k * (size_t)ar.vstride;
            x.S.step_off = k;
            x.S.stage = st;
            x.S.y = a->lam.p;
            x.S.acc = a->aacc.p;
            x.S.rin = st == 0 ? a->lam.p : (st == 2 ? a->q2.p : a->q1.p);
            x.S.out = st == 3 ? a->lam.p : (st == 1 ? a->q2.p : a->q1.p);
            x.Yf = a->trajm ? tslot(3 - st) : stg(k, 3 - st);
            for (const Group& g : im->groups)
                if ((rc = launch_adj(g, x, s))) {
                    hipGraph_t junk;
                    (void)hipStreamEndCapture(s, &junk);
                    return rc;
                }
        }
    }
    hipLaunchKernelGGL(tick, dim3(1), dim3(1), 0, s, im->step_ctr.p, -a->C);
    HIPCHK(hipStreamEndCapture(s, &a->gbwd));
    HIPCHK(hipGraphInstantiate(&a->xbwd, a->gbwd, nullptr, nullptr, 0));
    im->adj = a.release();
    return EH_OK;
}

// forward with checkpoints (called by batch_run when the adjoint is enabled)
int adj_forward(Batch* b) {
    BatchImpl* im = b->impl;
    AdjState* a = im->adj;
    double2* first = a->trajm ? a->traj.p : a->ckpt.p;
    HIPCHK(hipMemcpyAsync(first, im->Y.p, (size_t)a->total * sizeof(double2), hipMemcpyDeviceToDevice, im->stream));
    for (int c = 0; c < a->nseg; c++) HIPCHK(hipGraphLaunch(a->xfwd, im->stream));
    if (a->trajm) {   // final states back into Y (each instance ended at its own step count)
        hipLaunchKernelGGL(traj_final, dim3(b->n_inst), dim3(256), 0, im->stream, a->traj.p, im->Y.p, im->inst.p,
                           a->total);
        HIPCHK(hipGetLastError());
    }
    return EH_OK;
}

int batch_backward(Batch* b, const double* dP, int64_t n_dP) {
    if (!b || !b->impl || !b->impl->adj) return fail(EH_EINVAL, "backward: adjoint not enabled");
    BatchImpl* im = b->impl;
    AdjState* a = im->adj;
    if (!im->ran) return fail(EH_EINVAL, "backward before forward run");
    if (n_dP != (int64_t)im->P.n)
        return fail(EH_ESHAPE, "backward: dP has " + std::to_string(n_dP) + " entries, the batch's readouts " +
                                   std::to_string(im->P.n));
    HIPCHK(hipSetDevice(b->ctx->device));
    hipStream_t s = im->stream;
    if (im->P.n) HIPCHK(hipMemcpyAsync(a->gP.p, dP, im->P.n * sizeof(double2), hipMemcpyHostToDevice, s));
    const long long T = a->total;
    const int blocks = (int)std::min<long long>(std::max<long long>(1, (T + 255) / 256), 65536);
    hipLaunchKernelGGL(fill_zero, dim3(blocks), dim3(256), 0, s, a->lam.p, T);
    hipLaunchKernelGGL(fill_zero, dim3(256), dim3(256), 0, s, a->gacc.p, (long long)a->gacc.n);
    if (a->slots.n) hipLaunchKernelGGL(fill_zero, dim3(256), dim3(256), 0, s, a->slots.p, (long long)a->slots.n);
    hipLaunchKernelGGL(set_ctr, dim3(1), dim3(1), 0, s, im->step_ctr.p, (a->nseg - 1) * a->C);
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(im->ev0, s));
    for (int c = 0; c < a->nseg; c++) HIPCHK(hipGraphLaunch(a->xbwd, s));
    HIPCHK(hipEventRecord(im->ev1, s));
    HIPCHK(hipStreamSynchronize(s));
    a->h_gacc.resize(a->gacc.n);
    HIPCHK(hipMemcpy(a->h_gacc.data(), a->gacc.p, a->gacc.n * sizeof(double2), hipMemcpyDeviceToHost));
    for (int i = 0; i < b->n_inst; i++) {   // fold the copies into the first
        const InstDev& I = im->hinst[(size_t)i];
        const long long gs = gacc_size(I.N * I.N, I.d);
        double2* blk = a->h_gacc.data() + a->h_gacc_off[(size_t)i];
        for (int c = 1; c < a->ncopy; c++)
            for (long long e = 0; e < gs; e++) {
                blk[e].x += blk[c * gs + e].x;
                blk[e].y += blk[c * gs + e].y;
            }
    }
    if (a->slots.n) {   // and the per-tile blocks of the small-N tiles into their instance's
        a->h_slots.resize(a->slots.n);
        HIPCHK(hipMemcpy(a->h_slots.data(), a->slots.p, a->slots.n * sizeof(double2), hipMemcpyDeviceToHost));
        for (size_t t = 0; t < a->h_slot_off.size(); t++) {
            if (a->h_slot_off[t] < 0) continue;
            const int i = im->htile_inst[t];
            const InstDev& I = im->hinst[(size_t)i];
            const long long gs = gacc_size(I.N * I.N, I.d);
            double2* blk = a->h_gacc.data() + a->h_gacc_off[(size_t)i];
            const double2* src = a->h_slots.data() + a->h_slot_off[t];
            for (long long e = 0; e < gs; e++) {
                blk[e].x += src[e].x;
                blk[e].y += src[e].y;
            }
        }
    }
    a->have_grad = true;
    return EH_OK;
}

// gradients of instance i in the convention g = dL/dRe + i dL/dIm:
// dH (N*N), dX4 (4*N*N), dalpha, dalpha_t, dgamma (K*Km each), all complex interleaved
int batch_get_grad(Batch* b, int i, double* dH, double* dX4, double* dal, double* dalt, double* dga) {
    if (!b || !b->impl || !b->impl->adj || !b->impl->adj->have_grad) return fail(EH_EINVAL, "no gradient computed");
    if (i < 0 || i >= b->n_inst) return fail(EH_ERANGE, "instance index out of range");
    BatchImpl* im = b->impl;
    AdjState* a = im->adj;
    const InstDev& I = im->hinst[(size_t)i];
    const int N = I.N, NN = N * N, d = I.d;
    const double2* acc = a->h_gacc.data() + a->h_gacc_off[(size_t)i];
    auto C = [](double2 v) { return cplx(v.x, v.y); };
    const cplx mi(0, -1);
    auto put = [](double* o, size_t k, cplx v) {
        o[2 * k] = v.real();
        o[2 * k + 1] = v.imag();
    };
    // dL/dV = (-i Gt)^H  ->  dH[c][d] = conj(-i Gt[d][c])
    for (int c = 0; c < N; c++)
        for (int e = 0; e < N; e++) {
            if (dH) put(dH, (size_t)c * N + e, std::conj(mi * C(acc[e * N + c])));
            if (dX4) {
                // V = H - e1 X[0] - e2 X[1]^T - e3 X[2]^T
                put(dX4, (size_t)c * N + e, -std::conj(mi * C(acc[NN + e * N + c])));
                put(dX4, (size_t)NN + (size_t)e * N + c, -std::conj(mi * C(acc[2 * NN + e * N + c])));
                put(dX4, (size_t)2 * NN + (size_t)e * N + c, -std::conj(mi * C(acc[3 * NN + e * N + c])));
                put(dX4, (size_t)3 * NN + (size_t)c * N + e, C(acc[4 * NN + c * N + e]));
            }
        }
    const double2* Sg = acc + 5 * NN;
    const double2* U = Sg + d;
    const double2* Ap = U + d;
    const double2* Bp = Ap + d;
    for (int j = 0; j < I.K; j++)
        for (int k = 0; k < I.Km; k++) {
            const int p = j + k * I.K, o = j * I.Km + k;
            const cplx al = C(im->hal[(size_t)I.al_off + o]), alt = C(im->halt[(size_t)I.al_off + o]);
            if (dga) put(dga, (size_t)o, -std::conj(C(Sg[p])));
            const cplx Q = C(U[p]) - al * C(Ap[p]) - alt * C(Bp[p]);
            if (dal) put(dal, (size_t)o, Q.real() * al / (2.0 * std::norm(al)) + std::conj(C(Ap[p])));
            if (dalt) put(dalt, (size_t)o, std::conj(C(Bp[p])));
        }
    return EH_OK;
}

}  // namespace eh
