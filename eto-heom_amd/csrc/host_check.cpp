// host_check.cpp -- a CPU-only driver of the host-side sources of libetoheom.so
// (key parser, construct_Hal, dipole assembly, sample draws, hierarchy and coupling
// tables, time grid, batch plan) for the address / undefined-behaviour sanitizer build
// (`make sanitize`).  Every path named on the command line is parsed in both key modes;
// whatever parses is carried through every builder the library runs before device
// work.  One line per file and mode: "ok <path> <mode>" or "err <code> <path> <mode>".
// A sanitizer finding aborts the process with a nonzero status.
#include <cstdio>
#include <string>
#include <vector>

#include "host.hpp"

namespace eh {
static std::string g_err;
int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}
}  // namespace eh

using namespace eh;

static eh_system system_of(const Problem& k) {
    eh_system s = {};
    s.N = k.sys_size;
    s.K = k.K;
    s.Km = k.Km;
    s.L = k.L;
    s.site_offset = k.mode == EH_MODE_2DES ? 1 : 0;
    s.alpha = reinterpret_cast<const double*>(k.alpha.data());
    s.alpha_t = reinterpret_cast<const double*>(k.alpha_t.data());
    s.gamma = reinterpret_cast<const double*>(k.gamma.data());
    for (int q = 0; q < 3; q++) {
        s.pulse[q][0] = k.pulses[q].E0;
        s.pulse[q][1] = k.pulses[q].tau0;
        s.pulse[q][2] = k.pulses[q].taup;
        s.pulse[q][3] = k.pulses[q].w0;
    }
    s.t_start = k.t_start;
    s.t_end = k.t_end;
    s.step_size = k.step_size;
    s.print_step = k.print_step;
    return s;
}

// everything after a successful parse; returns the first error code or EH_OK
static int exercise(Problem& k) {
    const int N = k.sys_size, ss = k.single_size;
    if (N <= 0) return EH_EPARSE;
    std::vector<cplx> H((size_t)N * N);
    int rc = construct_hal(k, nullptr, H.data());
    if (rc) return rc;
    std::vector<double> dH((size_t)ss * ss, 0.0);
    for (int i = 0; i < ss; i++) dH[(size_t)i * ss + i] = 1e-3 * (i + 1);
    if ((rc = construct_hal(k, dH.data(), H.data()))) return rc;
    if (k.mode == EH_MODE_2DES) {
        double polar[12];
        polar_default(k, polar);
        const double c3[3] = {0.0, 0.6, 0.8}, a3[3] = {1.0, 0.0, 0.0};
        polar_rotate(c3, a3, polar);
        if ((rc = set_polar(k, polar))) return rc;
        if (k.n_sample > 0) {
            std::vector<cplx> Hs, Xs;
            if ((rc = draw_samples(k, std::min(k.n_sample, 3), Hs, Xs))) return rc;
        }
    }
    const eh_system sys = system_of(k);
    const int d = sys.K * sys.Km;
    if (d > 255 || sys.L < 0 || sys.L > 255) return EH_EINVAL;
    const int64_t M = hierarchy_size(d, sys.L);
    if (M < 0 || M > 20000 || (int64_t)N * N * M > (int64_t)1 << 24) return EH_OK;   // too big to enumerate here
    Hierarchy h;
    if ((rc = build_hierarchy(d, sys.L, h))) return rc;
    Tables tb;
    if ((rc = build_tables(h, sys, tb))) return rc;
    const int nv[3] = {1, 1, 1};
    Schedule s;
    if ((rc = build_schedule(sys, k.mode, nv, s))) return rc;
    if (k.mode == EH_MODE_2DES && (int)k.X.size() != 4 * N * N) return EH_OK;
    eh_instance in = {};
    in.sys = &sys;
    in.H = reinterpret_cast<const double*>(H.data());
    in.X4 = k.mode == EH_MODE_2DES ? reinterpret_cast<const double*>(k.X.data()) : nullptr;
    in.nv[0] = in.nv[1] = in.nv[2] = 1;
    in.mode = k.mode;
    const int sites[2] = {1, N};
    in.sites = sites;
    in.n_sites = N >= 2 ? 2 : 1;
    BatchPlan plan;
    const eh_instance two[2] = {in, in};
    return plan_batch(2, two, 32, plan);
}

int main(int argc, char** argv) {
    int ok = 0, err = 0;
    for (int i = 1; i < argc; i++) {
        for (int mode : {EH_MODE_2DES, EH_MODE_DYNAMICS}) {
            Problem k;
            int rc = parse_key(argv[i], mode, k);
            if (rc == EH_OK) rc = exercise(k);
            if (rc == EH_OK) {
                ok++;
                std::printf("ok %s %d\n", argv[i], mode);
            } else {
                err++;
                std::printf("err %d %s %d\n", rc, argv[i], mode);
            }
        }
    }
    std::printf("done %d ok %d err\n", ok, err);
    return 0;
}
