// device.hpp -- device-side structures and helpers shared by the forward
// (engine.hip) and adjoint (adjoint.hip) kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace eh {

struct InstDev {
    long long state_off;   // complex offset of ADO 0 in the state buffers
    long long p_off;       // complex offset in the readout buffer
    long long lk_off;      // offset of ADO 0's row in the (ADO, mode) neighbour tables
    int N, M;
    int ado_off;           // offset in the global ADO numbering (damping)
    int nsteps;
    int v_off;             // complex offset of this instance's N*N blocks (H, V, X'_q)
    int step_off;          // offset in the per-step tables
    int nread;             // readout operators
    int npoints;
    int r_off;             // complex offset of this instance's readout operators
    int d, K, Km;          // hierarchy dimension d = K*Km, baths, modes per bath
    int so;                // site offset: bath j couples state j + so
    int al_off;            // offset of alpha / alpha~ (K*Km complex)
    double h2, h, hb[4];
    long long col_off;     // complex offset of ADO 0 in the coupled-column copies (K*N per ADO)
    long long vp_off;      // offset of the zero-padded 64x64 real/imag planes of V (N > 32), else -1
    int hsym;              // H is real and symmetric (then so is V in the steps with no pulse on)
    int sp_off;            // row offset of this instance's sparsity pattern of V (sp_cnt, sp_idx)
};

struct StageArgs {
    const InstDev* inst;
    const int* tile_inst;
    const int* tile_ado0;
    const int* tile_nado;
    const int* step_ctr;
    int step_off;
    int stage;
    const uint8_t* vreal;    // per step: V(t_i) has no imaginary part
    const double2* V;        // per instance N*N
    const double* Vpad;      // per instance with N > 32: zero-padded 64x64 real, imag planes
    const int* nb_up;        // per (ADO, p): index of n + e_p or -1
    const int* nb_dn;        // per (ADO, p): index of n - e_p or -1
    const double* cx;        // per (ADO, p): sqrt((n_p + 1) |alpha_p|)
    const double* cy;        // per (ADO, p): sqrt(n_p / |alpha_p|)
    const double2* alpha;    // per instance K*Km
    const double2* alpha_t;
    const double2* damp;     // per global ADO: -sum_p n_p gamma_p
    double gsign;            // sign of the coupling weights (+1 forward, -1 adjoint)
    const double2* rin;      // stage input
    const double2* y;        // Y
    double2* acc;            // ACC
    double2* out;            // stage output
    // Coupled-column copies: per ADO the K columns s = j + so of its matrix, each
    // a contiguous N-vector ([ADO][j][row]).  A stage reads the neighbours' columns
    // from cin (the copy of rin) and writes the copy of its output to cout, so the
    // column gathers are contiguous instead of one cache line per element.  NULL:
    // columns are gathered from rin directly.
    const double2* cin;
    double2* cout;
    // Trajectory mode: rin, out and y are slots of a per-step array and move by
    // step * traj elements (0: fixed buffers)
    long long traj;
    // Fused steps (the forward graph): V is double-buffered by step parity (the odd
    // steps' copy vstride / vpstride elements on), stage 0's workgroup holding ADO 0
    // reads the instance out, and stage 3's builds V of the next step into the other
    // copy, so no per-step prologue launch is needed.  0: one V, built by the prologue.
    int fused;
    long long vstride, vpstride;
    const struct FusedDev* fz;   // what the fused tail reads (device memory: few kernel-argument registers)
    // 1: the small-N and sparse stages' state stores (out, ACC) carry the non-temporal hint.
    // Measured: 3 % faster on a group whose state is far larger than the last-level cache
    // (with the same L2 traffic and hit rate), 12 % slower on one that nearly fits it (the
    // next stage reads its input back from that cache); the sparse stage on the 91 GB chain
    // grid 1.6 % faster.  The MFMA kernels ignore it.
    int stream;
    // Sparsity pattern of V(t) over the whole run, per instance (row offset InstDev::sp_off):
    // row a holds sp_cnt[a] column indices sp_idx[a * kSpCap + r] (ascending, -1 beyond);
    // the pattern is symmetric, so row c's list is also column c's.
    const int* sp_cnt;
    const short* sp_idx;
};

constexpr int kSpCap = 16;   // entries per row of the sparsity pattern (the sparse kernel's limit)

struct FusedDev {
    double2* V;              // copy 0 of V and of Vpad (the odd steps' copies vstride / vpstride on)
    double* Vpad;
    long long vstride, vpstride;
    const double2* H;        // per instance N*N
    const double2* Xp;       // per instance 3*N*N: X[0], X[1]^T, X[2]^T
    const double2* field;    // per step 3 scalars
    const int* pidx;         // per step readout slot
    const double2* R;        // per instance nread*N*N readout operators
    double2* P;
};

// the stage's buffers at this step (trajectory mode)
__device__ __forceinline__ void at_step(StageArgs& A, int step) {
    if (A.traj) {
        const long long o = (long long)step * A.traj;
        A.rin += o;
        A.out += o;
        A.y += o;
    }
    if (A.fused && (step & 1)) {
        A.V += A.vstride;
        A.Vpad += A.vpstride;
    }
}

// Store of a state element, with the non-temporal hint when the kernel offers it (ST)
// and the stage streams (A.stream).
typedef double dvec2 __attribute__((ext_vector_type(2)));
template <bool ST>
__device__ __forceinline__ void st_state(const StageArgs& A, double2* p, double2 v) {
    if (ST && A.stream) {
        const dvec2 w = {v.x, v.y};
        __builtin_nontemporal_store(w, reinterpret_cast<dvec2*>(p));
    } else {
        *p = v;
    }
}

// A read of a table the kernel never writes (instance descriptors, tile lists, sparsity
// patterns, this step's V), through the constant address space: with a uniform index it
// is a scalar load.
template <class T>
__device__ __forceinline__ T ro(const T* p, long long i) {
#if defined(__HIP_DEVICE_COMPILE__)
    return ((const __attribute__((address_space(4))) T*)p)[i];
#else
    return p[i];   // host pass: never called
#endif
}

__device__ __forceinline__ double2 cmad(double2 acc, double2 a, double2 b) {
    acc.x = fma(a.x, b.x, fma(-a.y, b.y, acc.x));
    acc.y = fma(a.x, b.y, fma(a.y, b.x, acc.y));
    return acc;
}

// RK4 stage update of one element from its RHS value k (the reference's classic RK4,
// dynamics.cpp:386-417, in 7 reads + 5 writes of the state per step instead of 9 + 7).
// x is the element of this stage's input, y of Y, a of ACC (each read where used; no
// stage uses both, so a kernel may load one operand and pass it as both):
//   stage 0: out = x + h/2 k                     (x = y)
//   stage 1: ACC = (x + y)/3 + h/3 k,  out = y + h/2 k
//   stage 2: out = y + h k
//   stage 3: out (= Y) = ACC + x/3 + h/6 k
// ACC holds the classic accumulator y + h/6 k1 + h/3 k2 less y/3: stage 1 rebuilds the
// h/6 k1 term from its own input x = y + h/2 k1 (stage 0 writes no ACC), and stage 3's
// x/3 = y/3 + h/3 k3 restores y/3 and adds the h/3 k3 term from its own input x = y + h k3
// (stage 2 writes no ACC, stage 3 reads no Y).
// A kernel with no registers to spare passes two operands (TWO; x unused): y (x in stage
// 0) and a (x in stage 1, ACC + x/3 folded by the caller in stage 3).
// Element (row, col) of ADO m (instance-local index) also goes to the column copy.
template <bool ST = false, bool TWO = false>
__device__ __forceinline__ void rk_store(const StageArgs& A, const InstDev& I, size_t gi, double2 x, double2 y,
                                         double2 a, double kr, double ki, int m, int row, int col) {
    const int st = A.stage;
    if (TWO) x = st == 0 ? y : a;
    double2 o;
    if (st == 0) {
        o = make_double2(fma(I.h2, kr, x.x), fma(I.h2, ki, x.y));
    } else if (st == 1) {
        const double2 b = make_double2((x.x + y.x) * (1.0 / 3.0), (x.y + y.y) * (1.0 / 3.0));
        st_state<ST>(A, A.acc + gi, make_double2(fma(I.hb[1], kr, b.x), fma(I.hb[1], ki, b.y)));
        o = make_double2(fma(I.h2, kr, y.x), fma(I.h2, ki, y.y));
    } else if (st == 2) {
        o = make_double2(fma(I.h, kr, y.x), fma(I.h, ki, y.y));
    } else {
        const double2 b = TWO ? a : make_double2(fma(x.x, 1.0 / 3.0, a.x), fma(x.y, 1.0 / 3.0, a.y));
        o = make_double2(fma(I.hb[3], kr, b.x), fma(I.hb[3], ki, b.y));
    }
    st_state<ST>(A, A.out + gi, o);
    const int jc = col - I.so;
    if (A.cout && jc >= 0 && jc < I.K) A.cout[I.col_off + ((long long)m * I.K + jc) * I.N + row] = o;
}

// address of element (a, b) = (pos, s) [rc = 1] or (s, pos) [rc = 0] of neighbour mu
// for bath j: columns from the column copy when the stage has one
// U24: mu and the ADO's size N^2 (and K N) below 2^24, so the full-rate 24-bit multiply is
// exact; true for N > 16, where an instance's M N^2 < 2^31 elements bounds M below 2^23.
template <int NT = 0, bool U24 = false>
__device__ __forceinline__ const double2* nb_elem(const StageArgs& A, const InstDev& I, int mu, int j, int rc,
                                                  int pos) {
    // 32-bit offsets inside the instance (an instance holds < 2^32 elements), one 64-bit add
    const int N = NT > 0 ? NT : I.N, s = j + I.so;
    // the small-N kernels (NT > 0) keep no column copies (measured: no gain below N = 16)
    if (NT == 0 && rc == 1 && A.cin)
        return (A.cin + I.col_off) + (unsigned)((U24 ? __umul24(mu, I.K * N) : mu * I.K * N) + j * N + pos);
    return (A.rin + I.state_off) +
           (unsigned)((U24 ? __umul24(mu, N * N) : mu * N * N) + (rc == 0 ? s * N + pos : pos * N + s));
}

// the one state operand of rk_store besides the input for element gi (Y in stages 1-2,
// ACC in stage 3, the input value in stage 0): pass it as both y and a
__device__ __forceinline__ double2 rk_operand(const StageArgs& A, size_t gi, double2 rin_val) {
    const int st = A.stage;
    return st == 3 ? A.acc[gi] : ((st == 1 || st == 2) ? A.y[gi] : rin_val);
}

// Coupling gathers of one ADO into LDS.  For bath j (site s = j + so), the
// reference's one-hot products S_j rho_m and rho_m S_j (dynamics.cpp:143-179)
// touch only row s and column s, so
//   cp[(j*2 + 0)*N + b] = sum_k  i x rho_{n+e_p}[s][b] + i alpha y rho_{n-e_p}[s][b]      (row s)
//   cp[(j*2 + 1)*N + a] = sum_k -i x rho_{n+e_p}[a][s] - i alpha~ y rho_{n-e_p}[a][s]     (column s)
// with p = j + k K.  Every load is independent, so a thread issues all of its
// 2 Km gathers before it waits.
// BIG: a kernel of one ADO per tile (na = 1) and N > 16, which splits e by a float
// reciprocal (e < 2 K N <= 2^13: exact, as entry_split states) and addresses the neighbours
// with 24-bit products (nb_elem's U24)
template <bool BIG = false>
__device__ __forceinline__ void gather_couplings(const StageArgs& A, const InstDev& I, int n0, int na, double2* cp,
                                                 int tid, int nthr) {
    const int N = I.N, K = I.K, Km = I.Km, d = I.d;
    const int per = 2 * K * N;
    const int total = BIG ? per : per * na;
    const float inv2N = 1.0f / (float)(2 * N);
    for (int e = tid; e < total; e += nthr) {
        int g, j, rc, pos;
        if constexpr (BIG) {
            g = 0;
            j = (int)(((float)e + 0.5f) * inv2N);
            const int rem = e - j * 2 * N;
            rc = rem >= N ? 1 : 0;
            pos = rem - rc * N;
        } else {
            g = e / per;
            const int eg = e - g * per;
            j = eg / (2 * N);
            const int rem = eg - j * 2 * N;
            rc = rem / N;
            pos = rem - rc * N;
        }
        const long long row0 = I.lk_off + (long long)(n0 + g) * d;
        double2 acc = make_double2(0.0, 0.0);
        for (int k = 0; k < Km; ++k) {
            const long long q = row0 + j + k * K;
            const int mu = A.nb_up[q], md = A.nb_dn[q];
            const double2 xu = mu >= 0 ? *nb_elem<0, BIG>(A, I, mu, j, rc, pos) : make_double2(0.0, 0.0);
            const double2 xd = md >= 0 ? *nb_elem<0, BIG>(A, I, md, j, rc, pos) : make_double2(0.0, 0.0);
            const double x = A.cx[q], y = A.cy[q];
            const double2 al = rc == 0 ? A.alpha[I.al_off + j * Km + k] : A.alpha_t[I.al_off + j * Km + k];
            // up:   row  i x,  column -i x;   down: row  i alpha y,  column -i alpha~ y
            const double sg = rc == 0 ? A.gsign : -A.gsign;
            const double2 wu = make_double2(0.0, sg * x);
            const double2 wd = make_double2(-sg * al.y * y, sg * al.x * y);
            acc = cmad(acc, wu, xu);
            acc = cmad(acc, wd, xd);
        }
        cp[e] = acc;
    }
}

// Two-phase form of gather_couplings for small-N tiles.  Phase 1 (gather_idx)
// loads, for the first EC entries per thread and the first KMX modes, the
// neighbour indices and weights; phase 2 (gather_data) the neighbour rows and
// columns.  Every load is unconditional (clamped address, result selected), so the
// compiler's wait counts are exact: vector loads retire in order and a load behind
// a branch would make every later wait assume the shortest path.  Entries beyond
// EC per thread and modes beyond KMX are added by gather_rest.
template <int EC, int KMX>
struct NbIdx {
    int up[EC][KMX], dn[EC][KMX];
    double x[EC][KMX], y[EC][KMX];
    double2 al[EC][KMX];
};

// entry e of a tile's coupling rows/columns, e = ((g K + j) 2 + rc) N + pos, split into
// (ADO g, bath j, row/column rc, position pos).  NT > 0: N is a compile-time constant.
// The divisions by a runtime N and K go through float reciprocals: (x + 0.5) / n lies at
// least 0.5 / n from an integer, and for x < 2^21 the float product's error is below that
// (the N > 16 callers have one ADO per tile: e < 2 K N <= 2^13; g K + j < 2^16).
template <int NT>
__device__ __forceinline__ void entry_split(int e, int N, int K, float invN, float invK, int& g, int& j, int& rc,
                                            int& pos) {
    const int n = NT > 0 ? NT : N;
    const int q = NT > 0 ? e / NT : (int)(((float)e + 0.5f) * invN);
    pos = e - __mul24(q, n);   // small operands: full-rate 24-bit products
    rc = q & 1;
    const int gj = q >> 1;
    g = (int)(((float)gj + 0.5f) * invK);
    j = gj - __mul24(g, K);
}

template <int EC, int KMX, int NT = 0>
__device__ __forceinline__ void gather_idx(const StageArgs& A, const InstDev& I, int n0, int na, int tid, int nthr,
                                           NbIdx<EC, KMX>& X) {
    // raw values only: masking them here would make the thread wait for each load
    const int N = NT > 0 ? NT : I.N, K = I.K, Km = I.Km, d = I.d;
    const int per = 2 * K * N, total = per * na;
    const float invN = 1.0f / (float)N, invK = 1.0f / (float)K;
#pragma unroll
    for (int r = 0; r < EC; ++r) {
        const int e = tid + r * nthr;
        const int ec = e < total ? e : total - 1;
        int g, j, rc, pos;
        entry_split<NT>(ec, N, K, invN, invK, g, j, rc, pos);
        const unsigned q0 = (unsigned)((n0 + g) * d + j);   // row in the instance's tables
#pragma unroll
        for (int k = 0; k < KMX; ++k) {
            const int kc = k < Km ? k : Km - 1;
            const long long q = I.lk_off + (q0 + (unsigned)(kc * K));
            X.up[r][k] = A.nb_up[q];
            X.dn[r][k] = A.nb_dn[q];
            X.x[r][k] = A.cx[q];
            X.y[r][k] = A.cy[q];
            X.al[r][k] = rc == 0 ? A.alpha[I.al_off + j * Km + kc] : A.alpha_t[I.al_off + j * Km + kc];
        }
    }
}

template <int EC, int KMX, int NT = 0>
__device__ __forceinline__ void gather_data(const StageArgs& A, const InstDev& I, int na, int tid, int nthr,
                                            const NbIdx<EC, KMX>& X, double2* cp) {
    const int N = NT > 0 ? NT : I.N, K = I.K, Km = I.Km;
    const int per = 2 * K * N, total = per * na;
    const float invN = 1.0f / (float)N, invK = 1.0f / (float)K;
    double2 xu[EC][KMX], xd[EC][KMX];
    int rcs[EC];
#pragma unroll
    for (int r = 0; r < EC; ++r) {
        const int e = tid + r * nthr;
        const int ec = e < total ? e : total - 1;
        int g, j, rc, pos;
        entry_split<NT>(ec, N, K, invN, invK, g, j, rc, pos);
        rcs[r] = rc;
#pragma unroll
        for (int k = 0; k < KMX; ++k) {
            // NT == 0: the N > 16 kernels (24-bit neighbour offsets, as nb_elem states)
            const int mu = X.up[r][k], md = X.dn[r][k];
            xu[r][k] = *nb_elem<NT, NT == 0>(A, I, mu > 0 ? mu : 0, j, rc, pos);
            xd[r][k] = *nb_elem<NT, NT == 0>(A, I, md > 0 ? md : 0, j, rc, pos);
        }
    }
#pragma unroll
    for (int r = 0; r < EC; ++r) {
        const int e = tid + r * nthr;
        const double sg = rcs[r] == 0 ? A.gsign : -A.gsign;
        double2 acc = make_double2(0.0, 0.0);
#pragma unroll
        for (int k = 0; k < KMX; ++k) {
            // up: row i x, column -i x;  down: row i alpha y, column -i alpha~ y
            const bool ok = e < total && k < Km;
            const double x = ok && X.up[r][k] >= 0 ? X.x[r][k] : 0.0;
            const double y = ok && X.dn[r][k] >= 0 ? X.y[r][k] : 0.0;
            const double2 al = X.al[r][k];
            acc = cmad(acc, make_double2(0.0, sg * x), xu[r][k]);
            acc = cmad(acc, make_double2(-sg * al.y * y, sg * al.x * y), xd[r][k]);
        }
        if (e < total) cp[e] = acc;
    }
}

// what gather_idx / gather_data leave out: modes k >= KMX of the first EC entries
// per thread and every mode of the entries beyond; added into cp (same thread)
template <int EC, int KMX>
__device__ __forceinline__ void gather_rest(const StageArgs& A, const InstDev& I, int n0, int na, int tid, int nthr,
                                            double2* cp) {
    const int N = I.N, K = I.K, Km = I.Km, d = I.d;
    const int per = 2 * K * N, total = per * na;
    for (int e = tid; e < total; e += nthr) {
        const int r = (e - tid) / nthr;
        const int k0 = r < EC ? KMX : 0;
        if (k0 >= Km) continue;
        const int g = e / per, eg = e - g * per;
        const int j = eg / (2 * N), rem = eg - j * 2 * N;
        const int rc = rem / N, pos = rem - rc * N;
        const long long row0 = I.lk_off + (long long)(n0 + g) * d;
        const double sg = rc == 0 ? A.gsign : -A.gsign;
        double2 acc = r < EC ? cp[e] : make_double2(0.0, 0.0);
        for (int k = k0; k < Km; ++k) {
            const long long q = row0 + j + k * K;
            const int mu = A.nb_up[q], md = A.nb_dn[q];
            const double2 xu = mu >= 0 ? *nb_elem(A, I, mu, j, rc, pos) : make_double2(0.0, 0.0);
            const double2 xd = md >= 0 ? *nb_elem(A, I, md, j, rc, pos) : make_double2(0.0, 0.0);
            const double x = A.cx[q], y = A.cy[q];
            const double2 al = rc == 0 ? A.alpha[I.al_off + j * Km + k] : A.alpha_t[I.al_off + j * Km + k];
            acc = cmad(acc, make_double2(0.0, sg * x), xu);
            acc = cmad(acc, make_double2(-sg * al.y * y, sg * al.x * y), xd);
        }
        cp[e] = acc;
    }
}

// coupling term of element (row, col) from the gathered rows/columns
template <int NT = 0>
__device__ __forceinline__ double2 coupling_of(const InstDev& I, const double2* cp, int row, int col) {
    const int N = NT > 0 ? NT : I.N;
    double2 c = make_double2(0.0, 0.0);
    const int jr = row - I.so, jc = col - I.so;
    // 24-bit products (2 K + 1) N < 2^24: full-rate multiplies
    if (jr >= 0 && jr < I.K) {
        const double2 v = cp[__mul24(jr * 2 + 0, N) + col];
        c.x += v.x;
        c.y += v.y;
    }
    if (jc >= 0 && jc < I.K) {
        const double2 v = cp[__mul24(jc * 2 + 1, N) + row];
        c.x += v.x;
        c.y += v.y;
    }
    return c;
}

// rho_n[r][c] from its N x (N+1) LDS copy, 0 in the padding up to 64 (clamped read, selected)
__device__ __forceinline__ double2 lds_at(const double2* sR, int r, int c, int N, int LD) {
    const bool in = r < N && c < N;
    const double2 v = sR[(in ? r : 0) * LD + (in ? c : 0)];
    return in ? v : make_double2(0.0, 0.0);
}

// element (r, c) of an N x N complex matrix in global memory, 0 beyond N (clamped address,
// selected result: an unconditional load)
__device__ __forceinline__ double2 gl_at(const double2* M, int r, int c, int N) {
    const bool in = r < N && c < N;
    const double2 v = M[in ? r * N + c : 0];
    return in ? v : make_double2(0.0, 0.0);
}

typedef double dbl4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ dbl4 mfma4(double a, double b, dbl4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

struct PrologueArgs {
    const InstDev* inst;
    const int* ilist;        // instance of each workgroup (NULL: blockIdx.x)
    const int* step_ctr;
    int step_off;
    const double2* H;        // per instance N*N
    const double2* Xp;       // per instance 3*N*N: X[0], X[1]^T, X[2]^T
    const double2* field;    // per step 3 scalars
    const int* pidx;         // per step readout slot
    const double2* R;        // per instance nread*N*N readout operators
    const double2* y;
    double2* V;
    double* Vpad;            // zero-padded 64x64 real and imaginary planes of V (instances with N > 32)
    double2* P;
    int readout;             // 0: build V only (replays in the backward pass)
    long long traj;          // trajectory mode: y moves by step * traj elements
};

// V(t_i) = H - e_1 X[0] - e_2 X[1]^T - e_3 X[2]^T  (dynamics.cpp:348-363) with e the
// three field scalars of step i, by threads tid = 0..nthr-1; the zero-padded 64x64 real and
// imaginary planes as well when Vpad is not null (instances with N > 32).
// element ix of V = H - e_1 X[0] - e_2 X[1]^T - e_3 X[2]^T (X holds the three N*N blocks)
__device__ __forceinline__ double2 v_elem(const double2* H, const double2* X, int NN, double2 e0, double2 e1,
                                          double2 e2, int ix) {
    double2 v = H[ix];
    const double2 x0 = X[ix], x1 = X[NN + ix], x2 = X[2 * NN + ix];
    v.x -= x0.x * e0.x - x0.y * e0.y;
    v.y -= x0.x * e0.y + x0.y * e0.x;
    v.x -= x1.x * e1.x - x1.y * e1.y;
    v.y -= x1.x * e1.y + x1.y * e1.x;
    v.x -= x2.x * e2.x - x2.y * e2.y;
    v.y -= x2.x * e2.y + x2.y * e2.x;
    return v;
}

__device__ __forceinline__ void build_v(const InstDev& I, const double2* H, const double2* X, const double2* e,
                                        double2* V, double* Vpad, int tid, int nthr) {
    const int N = I.N, NN = N * N;
    const double2 e0 = e[0], e1 = e[1], e2 = e[2];
    auto vat = [&](int ix) { return v_elem(H, X, NN, e0, e1, e2, ix); };
    for (int q = tid; q < NN; q += nthr) V[q] = vat(q);
    if (Vpad) {
        double* vi = Vpad + 64 * 64;
        for (int q = tid; q < 64 * 64; q += nthr) {
            const int r = q >> 6, c = q & 63;
            const double2 v = (r < N && c < N) ? vat(r * N + c) : make_double2(0.0, 0.0);
            Vpad[q] = v.x;
            vi[q] = v.y;
        }
    }
}

// The readouts Tr(rho_0 R_r) before a step (dynamics.cpp:367-381, :246-253) into
// P[r * npoints + slot], by the 64 lanes of one wave (lane = 0..63, every lane present).
// One summation order wherever it runs, so every path gives the same bits.
__device__ __forceinline__ void readout_wave(const InstDev& I, const double2* rho0, const double2* R, int slot,
                                             double2* P, int lane = (int)threadIdx.x) {
    const int N = I.N, NN = N * N;
    for (int r = 0; r < I.nread; ++r) {
        const double2* Rr = R + (size_t)r * NN;
        double2 s = make_double2(0.0, 0.0);
        for (int q = lane; q < NN; q += 64) {
            const int j = q / N, k = q - j * N;
            s = cmad(s, rho0[q], Rr[k * N + j]);
        }
        for (int o = 32; o > 0; o >>= 1) {
            s.x += __shfl_xor(s.x, o, 64);
            s.y += __shfl_xor(s.y, o, 64);
        }
        if (lane == 0) P[(long long)r * I.npoints + slot] = s;
    }
}

// End of a fused stage, in the workgroup (or the group of waves) whose tile starts at
// the instance's ADO 0 (threads tid = 0..nthr-1 of it; tid 0..63 is one wave):
// stage 0 reads out rho_0 before the step (its input, which no stage-0 workgroup
// writes); stage 3 builds V of the next step into the copy the step after reads
// (that copy was last read by the previous step's stages, all finished).
__device__ __forceinline__ void fused_tail(const StageArgs& A, const InstDev& I, int step, int tid, int nthr) {
    const FusedDev& F = *A.fz;
    if (A.stage == 0) {
        const int slot = F.pidx[I.step_off + step];
        if (slot >= 0 && tid < 64) readout_wave(I, A.rin + I.state_off, F.R + I.r_off, slot, F.P + I.p_off, tid);
    } else if (A.stage == 3 && step + 1 < I.nsteps) {
        const long long o = ((step + 1) & 1) ? 1 : 0;
        build_v(I, F.H + I.v_off, F.Xp + 3 * (size_t)I.v_off, F.field + 3 * (size_t)(I.step_off + step + 1),
                F.V + o * F.vstride + I.v_off, I.vp_off >= 0 ? F.Vpad + o * F.vpstride + I.vp_off : nullptr, tid,
                nthr);
    }
}

// V(t_i) (build_v) and the readouts before step i (readout_wave) as a kernel of its
// own, one workgroup per instance: the per-launch and adjoint paths.
static __global__ __launch_bounds__(256) void prologue(PrologueArgs A) {
    const InstDev& I = A.inst[A.ilist ? A.ilist[blockIdx.x] : (int)blockIdx.x];
    const int step = *A.step_ctr + A.step_off;
    if (step >= I.nsteps) return;
    build_v(I, A.H + I.v_off, A.Xp + 3 * (size_t)I.v_off, A.field + 3 * (size_t)(I.step_off + step), A.V + I.v_off,
            I.vp_off >= 0 ? A.Vpad + I.vp_off : nullptr, threadIdx.x, blockDim.x);
    const int slot = A.pidx[I.step_off + step];
    if (slot < 0 || !A.readout || threadIdx.x >= 64) return;
    readout_wave(I, A.y + (long long)step * A.traj + I.state_off, A.R + I.r_off, slot, A.P + I.p_off);
}

static __global__ void tick(int* ctr, int by) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *ctr += by;
}

}  // namespace eh
