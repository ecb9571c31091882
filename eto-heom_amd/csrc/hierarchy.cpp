// hierarchy.cpp -- ADO enumeration, the neighbour/coefficient tables and the per-step schedule.
//
// The reference enumerates the multi-indices by DFS (etoheom-cpu/lib/utilize.cpp:3-18),
// which is the lexicographic order of {n in N^d : |n| <= L}; it then finds n +/- e_p
// by string edits and hash lookups on every RHS evaluation (dynamics.cpp:143-179).
// Here the lookups happen once, into (ADO, p) tables of neighbour indices and the
// reference's scaled coefficients, so the kernels only gather.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "host.hpp"

namespace eh {

int64_t hierarchy_size(int d, int L) {
    if (d < 0 || L < 0) return -1;
    // C(d+L, L) with overflow guard
    long double r = 1;
    for (int i = 1; i <= L; i++) r = r * (d + i) / i;
    if (r > 2e9L) return -1;
    return (int64_t)std::llround((double)r);
}

static void dfs(int d, int L, std::vector<uint8_t>& cur, int pos, int sum, std::vector<uint8_t>& out) {
    if (pos == d) {
        out.insert(out.end(), cur.begin(), cur.end());
        return;
    }
    for (int c = 0; c + sum <= L; c++) {
        cur[pos] = (uint8_t)c;
        dfs(d, L, cur, pos + 1, sum + c, out);
    }
}

int build_hierarchy(int d, int L, Hierarchy& h) {
    if (d < 0 || d > 255 || L < 0 || L > 255) return fail(EH_EINVAL, "hierarchy dimension / depth out of range");
    const int64_t M = hierarchy_size(d, L);
    if (M < 0) return fail(EH_ERANGE, "hierarchy too large");
    h.d = d;
    h.L = L;
    h.M = M;
    h.ado.clear();
    h.ado.reserve((size_t)(M * (d ? d : 1)));
    if (d == 0) {
        h.M = 1;
        return EH_OK;
    }
    std::vector<uint8_t> cur((size_t)d, 0);
    dfs(d, L, cur, 0, 0, h.ado);
    if ((int64_t)(h.ado.size() / d) != M) return fail(EH_ESHAPE, "hierarchy enumeration count mismatch");
    return EH_OK;
}

// lexicographic order -> binary search
int64_t Hierarchy::index_of(const uint8_t* key) const {
    if (d == 0) return 0;
    int64_t lo = 0, hi = M;
    while (lo < hi) {
        const int64_t mid = (lo + hi) / 2;
        const int c = std::memcmp(ado.data() + mid * d, key, (size_t)d);
        if (c == 0) return mid;
        if (c < 0)
            lo = mid + 1;
        else
            hi = mid;
    }
    return -1;
}

int build_tables(const Hierarchy& h, const eh_system& sys, Tables& t) {
    const int K = sys.K, Km = sys.Km, d = h.d;
    auto get = [](const double* a, int i) { return cplx(a[2 * i], a[2 * i + 1]); };
    for (int j = 0; j < K; j++)
        for (int k = 0; k < Km; k++)
            if (std::abs(get(sys.alpha, j * Km + k)) == 0.0)
                return fail(EH_EINVAL, "an ETO mode has |alpha| = 0 (the reference divides by it)");
    const size_t md = (size_t)h.M * d;
    t.up.assign(md, -1);
    t.dn.assign(md, -1);
    t.cx.assign(md, 0.0);
    t.cy.assign(md, 0.0);
    t.damp.assign((size_t)h.M, cplx(0, 0));
    std::vector<uint8_t> tmp((size_t)(d ? d : 1));
    for (int64_t i = 0; i < h.M; i++) {
        const uint8_t* n = h.ado.data() + i * d;
        int sum = 0;
        for (int p = 0; p < d; p++) sum += n[p];
        if (i != 0) {
            // -(sum_p n_p gamma_p), accumulated in the reference's j-then-k order (dynamics.cpp:130-138)
            cplx x(0, 0);
            for (int j = 0; j < K; j++)
                for (int k = 0; k < Km; k++) {
                    const int off = j + k * K;
                    if (n[off] == 0) continue;
                    x += cplx(-1, 0) * (get(sys.gamma, j * Km + k) * (double)n[off]);
                }
            t.damp[(size_t)i] = x;
        }
        for (int j = 0; j < K; j++)
            for (int k = 0; k < Km; k++) {
                const int p = j + k * K;
                const size_t q = (size_t)i * d + p;
                const double aabs = std::abs(get(sys.alpha, j * Km + k));
                if (n[p] < sys.L && sum < sys.L) {   // n + e_p exists (dynamics.cpp:146-149)
                    std::memcpy(tmp.data(), n, (size_t)d);
                    tmp[p] += 1;
                    const int64_t m = h.index_of(tmp.data());
                    if (m < 0) return fail(EH_ESHAPE, "missing up neighbour");
                    t.up[q] = (int32_t)m;
                    t.cx[q] = std::sqrt(((double)n[p] + 1) * aabs);
                }
                if (i != 0 && n[p] > 0) {             // n - e_p (dynamics.cpp:164-168)
                    std::memcpy(tmp.data(), n, (size_t)d);
                    tmp[p] -= 1;
                    const int64_t m = h.index_of(tmp.data());
                    if (m < 0) return fail(EH_ESHAPE, "missing down neighbour");
                    t.dn[q] = (int32_t)m;
                    t.cy[q] = std::sqrt((double)n[p] / aabs);
                }
            }
    }
    return EH_OK;
}

// Time grid and field of twoD_spectrum_solver / dynamics_solver.
int build_schedule(const eh_system& sys, int mode, const int nv[3], Schedule& s) {
    if (!(sys.step_size > 0)) return fail(EH_EINVAL, "TIME step_size must be positive");
    const float nstep = (sys.t_end - sys.t_start) / sys.step_size;   // float, dynamics.cpp:344
    const float pr = sys.print_step / sys.step_size;
    if (!(pr >= 1.0f && pr < 2147483647.0f)) return fail(EH_EINVAL, "print_step / step_size must be in [1, 2^31)");
    const int print_step = (int)pr;                                    // :346
    if (!(nstep < 1e8f)) return fail(EH_ERANGE, "too many time steps");
    s.nsteps = 0;
    for (int i = 0; i <= nstep; i++) s.nsteps++;
    s.e.assign((size_t)3 * s.nsteps, cplx(0, 0));
    s.pidx.assign((size_t)s.nsteps, -1);
    s.active.assign((size_t)s.nsteps, 0);
    s.npoints = 0;
    float t = sys.t_start;
    for (int i = 0; i < s.nsteps; i++) {
        if (mode == EH_MODE_2DES) {
            for (int q = 0; q < 3; q++) {
                const double E0 = sys.pulse[q][0], tau0 = sys.pulse[q][1], taup = sys.pulse[q][2], w0 = sys.pulse[q][3];
                if (!(nv[q] == 1 && std::fabs(t - tau0) <= taup * 2)) continue;
                // build_V_matrix / build_V_dagger_matrix, TD_hamiltonian.cpp:5-31
                const double E = E0 * std::exp(-2.77 * (t - tau0) * (t - tau0) / taup / taup);
                const double th = (q == 0) ? w0 * (t - tau0) : -w0 * (t - tau0);
                s.e[(size_t)3 * i + q] = cplx(E * std::cos(th), E * std::sin(th));
                s.active[(size_t)i] = 1;
            }
            if (t >= sys.pulse[2][1] && i % print_step == 0) s.pidx[(size_t)i] = s.npoints++;
        } else {
            if (i % print_step == 0) s.pidx[(size_t)i] = s.npoints++;
        }
        t += sys.step_size;   // float accumulation, dynamics.cpp:418
    }
    const double b[4] = {1.0 / 6, 2.0 / 6, 2.0 / 6, 1.0 / 6};
    s.h = sys.step_size;
    s.h2 = sys.step_size * 0.5;
    for (int q = 0; q < 4; q++) s.hb[q] = sys.step_size * b[q];
    return EH_OK;
}

}  // namespace eh
