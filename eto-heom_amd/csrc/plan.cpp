// plan.cpp -- host-only checks of a batch's instances and its device-memory plan.
//
// No HIP call here: eh_batch_plan answers "does this batch fit, and how many
// instances of it fit at once" before any device work, and batch_create runs the
// same per-instance checks.  The byte counts follow the allocations of
// batch_create (engine.hip) and batch_enable_adjoint (adjoint.hip).
#include <algorithm>
#include <map>
#include <vector>

#include "host.hpp"

namespace eh {

int check_instance(const eh_instance& q, InstPlan& p) {
    if (!q.sys || !q.H) return fail(EH_EINVAL, "instance without system or H");
    const eh_system& sys = *q.sys;
    const int N = sys.N;
    if (N <= 0 || N > kMaxN) return fail(EH_ESHAPE, "N must be in 1.." + std::to_string(kMaxN));
    if (q.mode != EH_MODE_2DES && q.mode != EH_MODE_DYNAMICS) return fail(EH_EINVAL, "unknown instance mode");
    if (q.mode == EH_MODE_2DES && !q.X4) return fail(EH_EINVAL, "2DES instance needs X4");
    if (q.mode == EH_MODE_DYNAMICS && (q.n_sites < 0 || (q.n_sites > 0 && !q.sites)))
        return fail(EH_EINVAL, "dynamics instance: bad sites");
    if (sys.K < 0 || sys.Km < 0 || sys.K + sys.site_offset > N) return fail(EH_ESHAPE, "bath sites exceed N");
    if ((sys.K > 0 && sys.Km > 0) && (!sys.alpha || !sys.alpha_t || !sys.gamma))
        return fail(EH_EINVAL, "bath modes without alpha / alpha_t / gamma");
    const int d = sys.K * sys.Km;
    if (d > 255 || sys.L < 0 || sys.L > 255) return fail(EH_EINVAL, "hierarchy dimension / depth out of range");
    const int64_t M = hierarchy_size(d, sys.L);
    if (M < 0) return fail(EH_ERANGE, "hierarchy too large");
    // the kernels address an instance's ADOs with 32-bit offsets (mu N^2 + a N + b)
    if (M * (int64_t)N * N >= (int64_t)1 << 31)
        return fail(EH_ERANGE, "instance state has " + std::to_string(M * (int64_t)N * N) +
                                   " elements; at most 2^31 - 1 per instance");
    Schedule s;
    const int rc = build_schedule(sys, q.mode, q.nv, s);
    if (rc) return rc;
    p.N = N;
    p.M = (int)(d == 0 ? 1 : M);
    p.d = d;
    p.nsteps = s.nsteps;
    p.npoints = s.npoints;
    p.nread = q.mode == EH_MODE_2DES ? 1 : q.n_sites;
    if (q.mode == EH_MODE_DYNAMICS && q.n_sites > N) return fail(EH_ERANGE, "number of output sites exceeds the system size");
    p.elems = (int64_t)p.M * N * N;
    p.col_elems = (int64_t)p.M * sys.K * N;
    return EH_OK;
}

int plan_batch(int n, const eh_instance* inst, int C, BatchPlan& out) {
    if (n <= 0 || !inst) return fail(EH_EINVAL, "eh_batch_plan: bad arguments");
    out = BatchPlan();
    const double z = 16.0;   // bytes per complex128
    double fixed = 0.0, vh_nn = 0.0;
    std::map<std::vector<double>, int> tables;   // hierarchy tables are shared by equal (d, L, alpha)
    int64_t n_big = 0;
    for (int i = 0; i < n; i++) {
        InstPlan p;
        const int rc = check_instance(inst[i], p);
        if (rc) return rc;
        const eh_system& sys = *inst[i].sys;
        out.total_ados += p.M;
        out.total_elems += p.elems;
        out.col_elems += p.col_elems;
        out.max_steps = std::max(out.max_steps, p.nsteps);
        const double NN = (double)p.N * p.N;
        // H, two V copies, X'_q, readout operators, field and step tables, readouts
        fixed += z * (NN * (6.0 + p.nread)) + (z * 3.0 + 5.0) * p.nsteps + z * (double)p.nread * p.npoints;
        vh_nn += NN;
        if (p.N > 32 && p.N <= 64) n_big++;
        std::vector<double> key = {(double)p.d, (double)sys.L, (double)sys.K, (double)sys.Km};
        if (p.d) key.insert(key.end(), sys.alpha, sys.alpha + 2 * p.d);
        if (tables.emplace(key, 1).second) fixed += (double)p.M * p.d * (4 + 4 + 8 + 8 + 1);
        fixed += z * p.M;   // damping
    }
    // zero-padded 64x64 real/imaginary V planes, two copies, per 32 < N <= 64 instance
    fixed += 2.0 * 2.0 * 64 * 64 * 8 * (double)n_big;
    const double T = (double)out.total_elems, Tc = (double)out.col_elems;
    // Y, ACC, S1, S2 and the three coupled-column copies
    out.forward_bytes = (int64_t)(z * (4.0 * T + 3.0 * Tc) + fixed);
    const int Ce = std::max(1, std::min(C > 0 ? C : 1, std::max(1, out.max_steps)));
    const int nseg = (out.max_steps + Ce - 1) / Ce;
    // lambda, its accumulator, two stage buffers; the adjoint tables; V^H of a segment
    const double common = z * 4.0 * T + fixed + z * Ce * vh_nn;
    out.adjoint_replay_bytes = (int64_t)(common + z * T * ((nseg + 1.0) + 4.0 * (Ce + 1.0) + 1.0));
    out.adjoint_traj_bytes = (int64_t)(common + z * T * (4.0 * out.max_steps + 1.0));
    return EH_OK;
}

}  // namespace eh
