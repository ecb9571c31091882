// host.hpp -- internal C++ declarations shared by the host-side sources of
// libetoheom.so (key parser, hierarchy builder, batch engine, C ABI).
#pragma once

#include <complex>
#include <cstdint>
#include <string>
#include <utility>
#include <vector>

#include "etoheom.h"

namespace eh {

using cplx = std::complex<double>;

constexpr double kHbar = 5.308e3;    // cm^-1 fs (etoheom-cpu/lib/constant.cpp:9)
constexpr double kBoltz = 0.695;     // cm^-1 / K (constant.cpp:8)
constexpr double kPi = 3.141592653589793238462643383279502884;   // constant.cpp:11

// thread-local error message behind eh_last_error()
int fail(int code, const std::string& msg);

struct Pulse {
    double E0 = 0, tau0 = 0, taup = 0, w0 = 0;
};

// The parsed key file: the data members of the reference `param` (param.h:38-82)
// that the propagation path reads.
struct Problem {
    int mode = EH_MODE_2DES;
    int single_size = 0, sys_size = 0;
    std::vector<std::pair<int, int>> d_map;   // index -> (a, b) for double-exciton states
    std::vector<cplx> Hal;                    // N*N
    std::vector<double> Disorder;             // N*N (sigma / h_bar)
    bool has_disorder = false, has_hal = false, has_dipole = false, has_polar = false;
    std::vector<cplx> Xc[3];                  // X_x, X_y, X_z
    std::vector<cplx> X;                      // 4*N*N
    std::vector<double> angle;                // radians
    int K = 0, Km = 0, L = 0;
    int n_sample = 0;
    unsigned long seed = 0;
    double beta = 0;
    std::vector<double> lambda;
    std::vector<cplx> alpha, alpha_t, gamma;  // K*Km, [j*Km + k]
    Pulse pulses[3];
    float t_start = 0, t_end = 0, step_size = 0, print_step = 0;
};

int parse_key(const std::string& path, int mode, Problem& out);
int construct_hal(const Problem& p, const double* dH, cplx* H_out);
void polar_default(const Problem& p, double* polar);
int set_polar(Problem& p, const double* polar);
void polar_rotate(const double* c3, const double* a3, double* polar);

// Hierarchy in the reference DFS order plus the CSR coupling table the kernels read.
struct Hierarchy {
    int d = 0, L = 0;
    int64_t M = 0;
    std::vector<uint8_t> ado;   // M*d
    int64_t index_of(const uint8_t* key) const;   // -1 if absent
};
int64_t hierarchy_size(int d, int L);
int build_hierarchy(int d, int L, Hierarchy& h);

// Coupling tables in (ADO, p) layout, p = j + k K (bath j, mode k):
//   up[n][p] = index of n + e_p or -1,  cx[n][p] = sqrt((n_p + 1) |alpha_p|)
//   dn[n][p] = index of n - e_p or -1,  cy[n][p] = sqrt(n_p / |alpha_p|)
// and damp[n] = -(sum_p n_p gamma_p) (zero for n = 0), which together carry the
// reference's scaled coefficients (dynamics.cpp:129-179).
struct Tables {
    std::vector<int32_t> up, dn;
    std::vector<double> cx, cy;
    std::vector<cplx> damp;
};
int build_tables(const Hierarchy& h, const eh_system& sys, Tables& t);

// Host-side per-step schedule of one propagation: float time grid, field scalars,
// readout slots (dynamics.cpp:344-418, TD_hamiltonian.cpp:5-31).
struct Schedule {
    int nsteps = 0;
    int npoints = 0;
    std::vector<cplx> e;         // 3 per step (0 when the pulse is off)
    std::vector<int32_t> pidx;   // readout slot before step i, or -1
    std::vector<uint8_t> active; // any pulse on at step i
    double h = 0, h2 = 0, hb[4] = {0, 0, 0, 0};
};
int build_schedule(const eh_system& sys, int mode, const int nv[3], Schedule& s);

void rng_stream(unsigned long seed, int count, uint32_t* out);
int draw_samples(const Problem& k, int n_sample, std::vector<cplx>& H, std::vector<cplx>& X);

// Largest system dimension the stage kernels take (the key parser's bound; an instance's
// state must also stay below 2^31 elements).
constexpr int kMaxN = 2048;

// One instance as checked on the host (plan.cpp): sizes and step counts.
struct InstPlan {
    int N = 0, M = 0, d = 0, nsteps = 0, npoints = 0, nread = 0;
    int64_t elems = 0, col_elems = 0;   // state elements M N^2, coupled-column elements M K N
};
// Every check batch_create makes on an instance before any device work.
int check_instance(const eh_instance& q, InstPlan& p);

// Device-memory plan of a batch (eh_batch_plan).
struct BatchPlan {
    int64_t total_ados = 0, total_elems = 0, col_elems = 0;
    int max_steps = 0;
    int64_t forward_bytes = 0, adjoint_replay_bytes = 0, adjoint_traj_bytes = 0;
};
int plan_batch(int n, const eh_instance* inst, int checkpoint_every, BatchPlan& out);

}  // namespace eh
