// keyfile.cpp -- the .key parser and the deterministic system assembly.
//
// parse_key follows param::param_2des (etoheom-cpu/lib/param.cpp:18-248) and
// param::param_dynamics (param.cpp:250-348) line for line in what it accepts:
// keyword lines are skipped when empty or starting with '#', data lines are
// read raw, unknown keywords are ignored, a block reads exactly the number of
// lines the reference reads (DISORDER reads sys_size rows), and a whitespace-only
// keyword line leaves the previous keyword in force as the reference's
// `iss >> key_word` does.  Where the reference would exit() or crash (PULSE count
// != 3, BATH before HAMILTONIAN, values past the matrix edge) an error code is
// returned instead.
#include <algorithm>
#include <cmath>
#include <fstream>
#include <sstream>

#include "host.hpp"

namespace eh {

namespace {

// system dimension of a 2DES key (ground + singles + doubles), in 64 bits
int64_t nsys_2des(int64_t n) { return n + (n - 1) * (n - 2) / 2; }

// Largest sys_size a key may declare: its N x N blocks (Hamiltonian, disorder, three
// dipole operators) are allocated as the blocks are read.
constexpr int64_t kMaxKeyN = 2048;
// Largest mode count of a BATH block (the hierarchy dimension K K_m is at most 255).
constexpr int kMaxModes = 255;

bool read_row(std::istream& f, std::string& line) { return static_cast<bool>(std::getline(f, line)); }

}  // namespace

int parse_key(const std::string& path, int mode, Problem& k) {
    if (mode != EH_MODE_2DES && mode != EH_MODE_DYNAMICS) return fail(EH_EINVAL, "mode must be EH_MODE_2DES or EH_MODE_DYNAMICS");
    std::ifstream f(path);
    if (!f) return fail(EH_EPARSE, "cannot open key file " + path);
    k = Problem();
    k.mode = mode;
    const bool twod = (mode == EH_MODE_2DES);
    std::string line, key_word;
    while (std::getline(f, line)) {
        if (line.empty() || line[0] == '#') continue;
        std::istringstream iss(line);
        iss >> key_word;   // on a blank-but-not-empty line the previous keyword stays (as in the reference)

        if (key_word == "HEOM") {
            read_row(f, line);
            std::istringstream v(line);
            v >> k.K >> k.L;
        } else if (key_word == "SIZE") {
            read_row(f, line);
            std::istringstream v(line);
            v >> k.single_size;
            if (k.single_size <= 0) return fail(EH_EPARSE, "SIZE must be positive");
            const int64_t ns = twod ? nsys_2des(k.single_size) : k.single_size;
            if (k.single_size > kMaxKeyN || ns > kMaxKeyN)
                return fail(EH_EPARSE, "SIZE gives a system dimension above " + std::to_string(kMaxKeyN));
            k.sys_size = (int)ns;
            k.d_map.assign(k.sys_size, {0, 0});
            if (twod) {
                int ct = k.single_size;
                for (int i = 1; i < k.single_size; i++)
                    for (int j = i; j < k.single_size; j++)
                        if (i != j) k.d_map[ct++] = {i, j};
            }
        } else if (twod && key_word == "DISORDER") {
            if (k.sys_size <= 0) return fail(EH_EPARSE, "DISORDER before SIZE");
            read_row(f, line);
            std::istringstream v(line);
            v >> k.n_sample >> k.seed;
            if (k.n_sample < 0) return fail(EH_EPARSE, "DISORDER n_sample must not be negative");
            const int N = k.sys_size;
            k.Disorder.assign((size_t)N * N, 0.0);
            for (int i = 0; i < N; i++) {
                if (!read_row(f, line)) break;
                std::istringstream row(line);
                double val;
                int j = 0;
                while (row >> val) {
                    if (j >= N) return fail(EH_EPARSE, "DISORDER row longer than sys_size");
                    k.Disorder[(size_t)i * N + j] = val / kHbar;
                    j++;
                }
            }
            k.has_disorder = true;
        } else if (key_word == "HAMILTONIAN") {
            if (k.sys_size <= 0) return fail(EH_EPARSE, "HAMILTONIAN before SIZE");
            const int N = k.sys_size;
            k.Hal.assign((size_t)N * N, cplx(0, 0));
            for (int i = 0; i < k.single_size; i++) {
                if (!read_row(f, line)) break;
                std::istringstream row(line);
                double val;
                int j = 0;
                while (row >> val) {
                    if (j >= N) return fail(EH_EPARSE, "HAMILTONIAN row longer than sys_size");
                    k.Hal[(size_t)i * N + j] = cplx(val / kHbar, 0.0);
                    j++;
                }
            }
            k.has_hal = true;
        } else if (twod && key_word == "DIPOLE") {
            if (k.sys_size <= 0) return fail(EH_EPARSE, "DIPOLE before SIZE");
            read_row(f, line);
            std::istringstream nd(line);
            int ndip = 0;
            nd >> ndip;
            const int N = k.sys_size;
            for (auto& m : k.Xc) m.assign((size_t)N * N, cplx(0, 0));
            for (int q = 0; q < ndip; q++) {
                if (!read_row(f, line)) return fail(EH_EPARSE, "DIPOLE block shorter than its count");
                std::istringstream ds(line);
                double x = 0, y = 0, z = 0;
                int i = 0, j = 0;
                ds >> i >> j >> x >> y >> z;
                if (i < 0 || j < 0 || i >= N || j >= N) return fail(EH_EPARSE, "DIPOLE index out of range");
                k.Xc[0][(size_t)i * N + j] = cplx(x, 0.0);
                k.Xc[1][(size_t)i * N + j] = cplx(y, 0.0);
                k.Xc[2][(size_t)i * N + j] = cplx(z, 0.0);
            }
            // double-exciton columns (param.cpp:117-146): X(i,(ab)) = d_bi X(0,a) + d_ai X(0,b)
            for (auto& X : k.Xc)
                for (int i = 0; i < N; i++)
                    for (int j = k.single_size; j < N; j++) {
                        const int a = k.d_map[j].first, b = k.d_map[j].second;
                        const double x1 = X[a].real(), x2 = X[b].real();   // row 0
                        const double x = (b == i ? 1.0 : 0.0) * x1 + (a == i ? 1.0 : 0.0) * x2;
                        X[(size_t)i * N + j] = cplx(x, 0.0);
                    }
            k.has_dipole = true;
        } else if (twod && key_word == "POLARIZATION") {
            read_row(f, line);
            std::istringstream ps(line);
            double ang[4] = {0, 0, 0, 0};
            ps >> ang[0] >> ang[1] >> ang[2] >> ang[3];
            for (int i = 0; i < 4; i++) k.angle.push_back(ang[i] * kPi / 180);
            k.X.assign((size_t)4 * k.sys_size * k.sys_size, cplx(0, 0));
            k.has_polar = true;
        } else if (key_word == "TEMPERATURE") {
            read_row(f, line);
            std::istringstream ts(line);
            double T = 0;
            ts >> T;
            k.beta = 1 / (T * kBoltz);
        } else if (key_word == "BATH") {
            if (!k.has_hal) return fail(EH_EPARSE, "BATH before HAMILTONIAN (the reference adds lambda to H)");
            read_row(f, line);
            std::istringstream bs(line);
            double l = 0;
            bs >> l;
            const int N = k.sys_size;
            const int so = twod ? 1 : 0;
            if (k.K < 0 || k.K + so > N) return fail(EH_EPARSE, "HEOM K exceeds the number of sites");
            for (int i = 0; i < k.K; i++) k.lambda.push_back(l);
            for (int i = 0; i < k.K; i++) {
                if (i + so >= N) return fail(EH_EPARSE, "HEOM K exceeds the number of sites");
                cplx& h = k.Hal[(size_t)(i + so) * N + (i + so)];
                h = cplx(h.real() + k.lambda[i], 0.0);
            }
            read_row(f, line);
            std::istringstream ms(line);
            int n_modes = 0;
            ms >> n_modes;
            if (n_modes < 0 || n_modes > kMaxModes)
                return fail(EH_EPARSE, "BATH mode count must be in 0.." + std::to_string(kMaxModes));
            std::vector<cplx> al, alt, ga;
            for (int i = 0; i < n_modes; i++) {
                if (!read_row(f, line)) return fail(EH_EPARSE, "BATH block shorter than its mode count");
                std::istringstream md(line);
                double a = 0, b = 0, g = 0, w = 0;
                md >> a >> b >> g >> w;
                al.emplace_back(a, b);
                alt.emplace_back(a, -b);
                ga.emplace_back(g, w);
            }
            // the reference appends one mode list per bath (param.cpp:212-216) and
            // reads K_m from the first; the first BATH block therefore wins
            if (k.Km == 0 || k.alpha.empty()) {
                k.alpha.clear();
                k.alpha_t.clear();
                k.gamma.clear();
                for (int i = 0; i < k.K; i++) {
                    k.alpha.insert(k.alpha.end(), al.begin(), al.end());
                    k.alpha_t.insert(k.alpha_t.end(), alt.begin(), alt.end());
                    k.gamma.insert(k.gamma.end(), ga.begin(), ga.end());
                }
                k.Km = n_modes;
            }
        } else if (twod && key_word == "PULSE") {
            read_row(f, line);
            std::istringstream ns(line);
            int n = 0;
            ns >> n;
            if (n != 3) return fail(EH_EPULSE, "Only support 3 pulses simulation.");
            for (int i = 0; i < 3; i++) {
                read_row(f, line);
                std::istringstream ps(line);
                ps >> k.pulses[i].E0 >> k.pulses[i].tau0 >> k.pulses[i].taup >> k.pulses[i].w0;
                k.pulses[i].w0 = k.pulses[i].w0 / kHbar;
                k.pulses[i].E0 = k.pulses[i].E0 / kHbar;
            }
        } else if (key_word == "TIME") {
            read_row(f, line);
            std::istringstream ts(line);
            ts >> k.t_start >> k.t_end >> k.step_size >> k.print_step;
        }
    }
    if (k.sys_size <= 0) return fail(EH_EPARSE, "key file has no SIZE block: " + path);
    if (!k.has_hal) k.Hal.assign((size_t)k.sys_size * k.sys_size, cplx(0, 0));
    if (k.Disorder.empty()) k.Disorder.assign((size_t)k.sys_size * k.sys_size, 0.0);
    return EH_OK;
}

// construct_Hal, etoheom-cpu/lib/dynamics.cpp:3-51.  dH holds the disorder draws
// (single_size^2, row-major) that the reference takes from gsl_ran_gaussian.
int construct_hal(const Problem& k, const double* dH, cplx* H) {
    const int N = k.sys_size, ss = k.single_size;
    for (int i = 0; i < N * N; i++) H[i] = k.Hal[i];
    if (dH)
        for (int i = 0; i < ss; i++)
            for (int j = 0; j < ss; j++) H[i * N + j] += cplx(dH[i * ss + j], 0.0);
    // ascending sort of the one-exciton diagonal; slot 0 holds 0.0 and takes part (:24-27)
    std::vector<double> Htmp(ss, 0.0);
    for (int i = 1; i < ss; i++) Htmp[i] = H[i * N + i].real();
    std::sort(Htmp.begin(), Htmp.end());
    for (int i = 1; i < ss; i++) H[i * N + i] = cplx(Htmp[i], 0.0);
    auto delta = [](int a, int b) { return a == b ? 1.0 : 0.0; };
    for (int i = ss; i < N; i++)
        for (int j = ss; j < N; j++) {
            const int a = k.d_map[i].first, b = k.d_map[i].second;
            if (i == j) {
                H[i * N + i] = H[a * N + a] + H[b * N + b];
            } else {
                const int c = k.d_map[j].first, d = k.d_map[j].second;
                const double x1 = delta(a, c) * (1 - delta(b, d)) * H[b * N + d].real();
                const double x2 = delta(a, d) * (1 - delta(b, c)) * H[b * N + c].real();
                const double x3 = delta(b, c) * (1 - delta(a, d)) * H[a * N + d].real();
                const double x4 = delta(b, d) * (1 - delta(a, c)) * H[a * N + c].real();
                H[i * N + j] = cplx(x1 + x2 + x3 + x4, 0.0);
            }
        }
    return EH_OK;
}

// polar_mat_set, polar.cpp:17-29: rows (0, sin theta_i, cos theta_i)
void polar_default(const Problem& k, double* polar) {
    for (int i = 0; i < 4; i++) {
        const double a = i < (int)k.angle.size() ? k.angle[i] : 0.0;
        polar[0 * 4 + i] = 0.0;
        polar[1 * 4 + i] = std::sin(a);
        polar[2 * 4 + i] = std::cos(a);
    }
}

// polar_mat_ranrot, polar.cpp:31-72, with the two random directions given
void polar_rotate(const double* c_in, const double* a_in, double* polar) {
    double c[3] = {c_in[0], c_in[1], c_in[2]}, a[3] = {a_in[0], a_in[1], a_in[2]}, b[3];
    auto norm = [](double* v) {
        const double s = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        v[0] /= s;
        v[1] /= s;
        v[2] /= s;
    };
    norm(c);
    const double pr = a[0] * c[0] + a[1] * c[1] + a[2] * c[2];
    for (int i = 0; i < 3; i++) a[i] = a[i] - pr * c[i];
    norm(a);
    b[0] = c[1] * a[2] - c[2] * a[1];
    b[1] = c[2] * a[0] - c[0] * a[2];
    b[2] = c[0] * a[1] - c[1] * a[0];
    const double* rot[3] = {a, b, c};
    double out[12];
    for (int r = 0; r < 3; r++)
        for (int i = 0; i < 4; i++) {
            double s = 0;
            for (int q = 0; q < 3; q++) s += rot[r][q] * polar[q * 4 + i];
            out[r * 4 + i] = s;
        }
    for (int i = 0; i < 12; i++) polar[i] = out[i];
}

// compute_pulse_interaction, polar.cpp:74-101
int set_polar(Problem& k, const double* polar) {
    if (!k.has_dipole) return fail(EH_EPARSE, "key file has no DIPOLE block");
    const int N = k.sys_size;
    k.X.assign((size_t)4 * N * N, cplx(0, 0));
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 3; j++) {
            const cplx E(polar[j * 4 + i], 0.0);
            for (int e = 0; e < N * N; e++) k.X[(size_t)i * N * N + e] += k.Xc[j][e] * E;
        }
    k.has_polar = true;
    return EH_OK;
}

}  // namespace eh
