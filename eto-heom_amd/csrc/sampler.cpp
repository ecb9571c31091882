// sampler.cpp -- the n_sample loop of the reference drivers.
//
// CPU_2DES (etoheom-cpu/src/CPU_2DES.cpp:14-96) seeds GSL's default generator
// with the key's seed and, per sample, draws the static disorder in
// construct_Hal (dynamics.cpp:13-20, gsl_ran_gaussian) and two random
// directions in polar_mat_ranrot (polar.cpp:45,48, gsl_ran_dir_3d).  GSL 2.8 is
// not vendored by the reference; its generator and the two variates are
// restated here from the published methods GSL documents for them:
//   * gsl_rng_mt19937: Matsumoto & Nishimura's MT19937 with the 2002
//     init_genrand seeding (seed 0 maps to 4357), uniform = u32 / 2^32,
//     uniform_pos rejects 0;
//   * gsl_ran_gaussian: Marsaglia's polar method (Knuth TAOCP 2, 3.4.1 Alg. P)
//     on uniform_pos, returning the y variate: sigma * y * sqrt(-2 ln r2 / r2);
//   * gsl_ran_dir_3d: Knop's method (CACM 13 (1970) 326; Knuth 2, 3.4.1):
//     (x, y) uniform in the unit disc, s = x^2 + y^2, z = 2s - 1,
//     (x, y) scaled by 2 sqrt(1 - s).
// The MT19937 stream is checked against std::mt19937; the variates' bit
// patterns against GSL cannot be checked here (no GSL in the image).
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "host.hpp"

namespace eh {

namespace {

class Mt19937 {
    uint32_t mt_[624];
    int mti_ = 624;

public:
    explicit Mt19937(unsigned long seed) {
        uint32_t s = (uint32_t)(seed & 0xffffffffUL);
        if (s == 0) s = 4357;
        mt_[0] = s;
        for (int i = 1; i < 624; i++) mt_[i] = 1812433253U * (mt_[i - 1] ^ (mt_[i - 1] >> 30)) + (uint32_t)i;
        mti_ = 624;
    }
    uint32_t next() {
        static const uint32_t mag01[2] = {0x0U, 0x9908b0dfU};
        if (mti_ >= 624) {
            int k = 0;
            for (; k < 624 - 397; k++) {
                const uint32_t y = (mt_[k] & 0x80000000U) | (mt_[k + 1] & 0x7fffffffU);
                mt_[k] = mt_[k + 397] ^ (y >> 1) ^ mag01[y & 1U];
            }
            for (; k < 623; k++) {
                const uint32_t y = (mt_[k] & 0x80000000U) | (mt_[k + 1] & 0x7fffffffU);
                mt_[k] = mt_[k + (397 - 624)] ^ (y >> 1) ^ mag01[y & 1U];
            }
            const uint32_t y = (mt_[623] & 0x80000000U) | (mt_[0] & 0x7fffffffU);
            mt_[623] = mt_[396] ^ (y >> 1) ^ mag01[y & 1U];
            mti_ = 0;
        }
        uint32_t y = mt_[mti_++];
        y ^= (y >> 11);
        y ^= (y << 7) & 0x9d2c5680U;
        y ^= (y << 15) & 0xefc60000U;
        y ^= (y >> 18);
        return y;
    }
    double uniform() { return next() / 4294967296.0; }
    double uniform_pos() {
        double x;
        do {
            x = uniform();
        } while (x == 0);
        return x;
    }
    double gaussian(double sigma) {
        double x, y, r2;
        do {
            x = -1 + 2 * uniform_pos();
            y = -1 + 2 * uniform_pos();
            r2 = x * x + y * y;
        } while (r2 > 1.0 || r2 == 0);
        return sigma * y * std::sqrt(-2.0 * std::log(r2) / r2);
    }
    void dir_3d(double* v) {
        double s;
        do {
            v[0] = -1.0 + 2.0 * uniform();
            v[1] = -1.0 + 2.0 * uniform();
            s = v[0] * v[0] + v[1] * v[1];
        } while (s > 1.0);
        v[2] = -1.0 + 2.0 * s;
        const double a = 2.0 * std::sqrt(1.0 - s);
        v[0] *= a;
        v[1] *= a;
    }
};

}  // namespace

// first `count` raw 32-bit outputs of the generator (for tests)
void rng_stream(unsigned long seed, int count, uint32_t* out) {
    Mt19937 g(seed);
    for (int i = 0; i < count; i++) out[i] = g.next();
}

// per sample: realised H (N*N complex) and X (4*N*N complex), in the draw order of
// CPU_2DES.cpp:50-60 (construct_Hal, then polar_mat_set + polar_mat_ranrot)
int draw_samples(const Problem& k, int n_sample, std::vector<cplx>& H, std::vector<cplx>& X) {
    if (k.mode != EH_MODE_2DES) return fail(EH_EINVAL, "sampling needs a 2DES problem");
    if (!k.has_dipole) return fail(EH_EPARSE, "key file has no DIPOLE block");
    if (k.angle.size() < 4) return fail(EH_EPARSE, "key file has no POLARIZATION block");
    const int N = k.sys_size, ss = k.single_size;
    Mt19937 g(k.seed);
    H.assign((size_t)n_sample * N * N, cplx(0, 0));
    X.assign((size_t)n_sample * 4 * N * N, cplx(0, 0));
    std::vector<double> dH((size_t)ss * ss);
    Problem w = k;   // set_polar writes w.X
    for (int s = 0; s < n_sample; s++) {
        for (int i = 0; i < ss; i++)
            for (int j = 0; j < ss; j++) {
                const double sigma = k.Disorder[(size_t)i * N + j];
                dH[(size_t)i * ss + j] = std::fabs(sigma) < 1e-10 ? 0.0 : g.gaussian(sigma);
            }
        int rc = construct_hal(k, dH.data(), H.data() + (size_t)s * N * N);
        if (rc) return rc;
        double polar[12], c[3], a[3];
        polar_default(k, polar);
        g.dir_3d(c);
        g.dir_3d(a);
        polar_rotate(c, a, polar);
        rc = set_polar(w, polar);
        if (rc) return rc;
        std::memcpy(X.data() + (size_t)s * 4 * N * N, w.X.data(), sizeof(cplx) * 4 * N * N);
    }
    return EH_OK;
}

}  // namespace eh
