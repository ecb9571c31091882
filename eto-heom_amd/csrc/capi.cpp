// capi.cpp -- the extern "C" entry points declared in include/etoheom.h.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <new>
#include <string>

#include "engine.hpp"

struct eh_problem {
    eh::Problem p;
    eh_system sys;   // view into p
};

struct eh_batch {
    eh::Batch* b = nullptr;
};

namespace eh {

static thread_local std::string g_err;

int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

static void fill_system(eh_problem* pr) {
    Problem& k = pr->p;
    eh_system& s = pr->sys;
    std::memset(&s, 0, sizeof(s));
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
}

}  // namespace eh

using namespace eh;

extern "C" {

const char* eh_last_error(void) { return g_err.c_str(); }

int eh_problem_from_key(const char* path, int mode, eh_problem** out) {
    if (!path || !out) return fail(EH_EINVAL, "eh_problem_from_key: NULL argument");
    std::unique_ptr<eh_problem> pr(new (std::nothrow) eh_problem());
    if (!pr) return fail(EH_ENOMEM, "out of memory");
    int rc = parse_key(path, mode, pr->p);
    if (rc) return rc;
    fill_system(pr.get());
    *out = pr.release();
    return EH_OK;
}

int eh_problem_destroy(eh_problem* p) {
    delete p;
    return EH_OK;
}

int eh_problem_dims(const eh_problem* pr, int64_t* o) {
    if (!pr || !o) return fail(EH_EINVAL, "eh_problem_dims: NULL argument");
    const Problem& k = pr->p;
    o[0] = k.mode;
    o[1] = k.single_size;
    o[2] = k.sys_size;
    o[3] = k.K;
    o[4] = k.Km;
    o[5] = k.L;
    o[6] = hierarchy_size(k.K * k.Km, k.L);
    o[7] = k.n_sample;
    o[8] = (int64_t)k.seed;
    return EH_OK;
}

int eh_problem_system(const eh_problem* pr, eh_system* sys) {
    if (!pr || !sys) return fail(EH_EINVAL, "eh_problem_system: NULL argument");
    *sys = pr->sys;
    return EH_OK;
}

int64_t eh_problem_get(const eh_problem* pr, const char* name, double* out, int64_t cap) {
    if (!pr || !name) return fail(EH_EINVAL, "eh_problem_get: NULL argument");
    const Problem& k = pr->p;
    std::vector<double> v;
    auto putc = [&](const std::vector<cplx>& m) {
        for (const cplx& c : m) {
            v.push_back(c.real());
            v.push_back(c.imag());
        }
    };
    const std::string n(name);
    if (n == "Hal") putc(k.Hal);
    else if (n == "Disorder") v = k.Disorder;
    else if (n == "X_x") putc(k.Xc[0]);
    else if (n == "X_y") putc(k.Xc[1]);
    else if (n == "X_z") putc(k.Xc[2]);
    else if (n == "X") putc(k.X);
    else if (n == "angle") v = k.angle;
    else if (n == "lambda") v = k.lambda;
    else if (n == "beta") v.push_back(k.beta);
    else if (n == "alpha") putc(k.alpha);
    else if (n == "alpha_t") putc(k.alpha_t);
    else if (n == "gamma") putc(k.gamma);
    else return fail(EH_EINVAL, "unknown problem array: " + n);
    if (!out) return (int64_t)v.size();
    if ((int64_t)v.size() > cap) return fail(EH_ERANGE, "output capacity too small for " + n);
    std::memcpy(out, v.data(), v.size() * sizeof(double));
    return (int64_t)v.size();
}

int eh_problem_polar_default(const eh_problem* pr, double* polar) {
    if (!pr || !polar) return fail(EH_EINVAL, "eh_problem_polar_default: NULL argument");
    polar_default(pr->p, polar);
    return EH_OK;
}

int eh_problem_set_polar(eh_problem* pr, const double* polar) {
    if (!pr || !polar) return fail(EH_EINVAL, "eh_problem_set_polar: NULL argument");
    return set_polar(pr->p, polar);
}

int eh_polar_rotate(const double* c3, const double* a3, double* polar) {
    if (!c3 || !a3 || !polar) return fail(EH_EINVAL, "eh_polar_rotate: NULL argument");
    polar_rotate(c3, a3, polar);
    return EH_OK;
}

int eh_construct_hal(const eh_problem* pr, const double* dH, double* H_out) {
    if (!pr || !H_out) return fail(EH_EINVAL, "eh_construct_hal: NULL argument");
    if (pr->p.mode != EH_MODE_2DES) return fail(EH_EINVAL, "construct_Hal needs a 2DES problem");
    return construct_hal(pr->p, dH, reinterpret_cast<cplx*>(H_out));
}

int eh_problem_samples(const eh_problem* pr, int n_sample, double* H_out, double* X_out) {
    if (!pr || n_sample < 0 || (n_sample > 0 && (!H_out || !X_out))) return fail(EH_EINVAL, "eh_problem_samples: bad arguments");
    std::vector<cplx> H, X;
    int rc = draw_samples(pr->p, n_sample, H, X);
    if (rc) return rc;
    std::memcpy(H_out, H.data(), H.size() * sizeof(cplx));
    std::memcpy(X_out, X.data(), X.size() * sizeof(cplx));
    return EH_OK;
}

int eh_rng_stream(unsigned long seed, int count, uint32_t* out) {
    if (count < 0 || (count > 0 && !out)) return fail(EH_EINVAL, "eh_rng_stream: bad arguments");
    rng_stream(seed, count, out);
    return EH_OK;
}

int64_t eh_hierarchy_size(int d, int L) {
    const int64_t m = hierarchy_size(d, L);
    return m < 0 ? fail(EH_ERANGE, "hierarchy too large") : m;
}

int eh_hierarchy_build(int d, int L, uint8_t* ado_out, int32_t* up_out, int32_t* dn_out) {
    Hierarchy h;
    int rc = build_hierarchy(d, L, h);
    if (rc) return rc;
    if (ado_out && d > 0) std::memcpy(ado_out, h.ado.data(), h.ado.size());
    std::vector<uint8_t> tmp((size_t)(d ? d : 1));
    for (int64_t i = 0; i < h.M; i++) {
        const uint8_t* n = h.ado.data() + i * d;
        int sum = 0;
        for (int p = 0; p < d; p++) sum += n[p];
        for (int p = 0; p < d; p++) {
            std::memcpy(tmp.data(), n, (size_t)d);
            if (up_out) {
                int64_t m = -1;
                if (sum < L) {
                    tmp[p] += 1;
                    m = h.index_of(tmp.data());
                    tmp[p] -= 1;
                }
                up_out[i * d + p] = (int32_t)m;
            }
            if (dn_out) {
                int64_t m = -1;
                if (n[p] > 0) {
                    tmp[p] -= 1;
                    m = h.index_of(tmp.data());
                }
                dn_out[i * d + p] = (int32_t)m;
            }
        }
    }
    return EH_OK;
}

int eh_ctx_create(int device, eh_ctx** out) {
    if (!out) return fail(EH_EINVAL, "eh_ctx_create: NULL argument");
    std::unique_ptr<eh_ctx> c(new (std::nothrow) eh_ctx());
    if (!c) return fail(EH_ENOMEM, "out of memory");
    int rc = ctx_init(device, c.get());
    if (rc) return rc;
    *out = c.release();
    return EH_OK;
}

int eh_ctx_destroy(eh_ctx* ctx) {
    delete ctx;
    return EH_OK;
}

int eh_ctx_info(const eh_ctx* ctx, char* name, int cap, int* n_cu, int* lds) {
    if (!ctx) return fail(EH_EINVAL, "eh_ctx_info: NULL context");
    if (name && cap > 0) {
        std::strncpy(name, ctx->name, (size_t)cap - 1);
        name[cap - 1] = 0;
    }
    if (n_cu) *n_cu = ctx->n_cu;
    if (lds) *lds = ctx->lds_per_block;
    return EH_OK;
}

int eh_ctx_mem_info(eh_ctx* ctx, int64_t* free_bytes, int64_t* total_bytes) {
    if (!ctx) return fail(EH_EINVAL, "eh_ctx_mem_info: NULL context");
    return ctx_mem_info(ctx, free_bytes, total_bytes);
}

int eh_batch_create(eh_ctx* ctx, int n, const eh_instance* inst, eh_batch** out) {
    if (!out) return fail(EH_EINVAL, "eh_batch_create: NULL argument");
    Batch* b = nullptr;
    int rc = batch_create(ctx, n, inst, &b);
    if (rc) return rc;
    *out = new eh_batch{b};
    return EH_OK;
}

int eh_batch_destroy(eh_batch* b) {
    if (b) batch_destroy(b->b);
    delete b;
    return EH_OK;
}

int eh_batch_run(eh_batch* b) { return b ? batch_run(b->b) : fail(EH_EINVAL, "NULL batch"); }
int eh_batch_sync(eh_batch* b) { return b ? batch_sync(b->b) : fail(EH_EINVAL, "NULL batch"); }
int eh_batch_elapsed_ms(eh_batch* b, double* ms) {
    return (b && ms) ? batch_elapsed_ms(b->b, ms) : fail(EH_EINVAL, "NULL argument");
}

int eh_batch_get_output(eh_batch* bb, int i, double* out, int64_t cap, int64_t* n_points) {
    if (!bb || !bb->b) return fail(EH_EINVAL, "NULL batch");
    Batch* b = bb->b;
    if (i < 0 || i >= b->n_inst) return fail(EH_ERANGE, "instance index out of range");
    if (!b->synced) {   // fails with EH_EINVAL when the batch has not run
        int rc = batch_sync(b);
        if (rc) return rc;
    }
    const Batch::Info& f = b->info[(size_t)i];
    if (n_points) *n_points = f.npoints;
    if (!out) return EH_OK;
    if (f.npoints > cap) return fail(EH_ERANGE, "output capacity too small");
    bool finite = true;
    if (f.mode == EH_MODE_2DES) {
        for (int t = 0; t < f.npoints; t++) {
            const cplx c = b->P_host[(size_t)(f.p_off + t)];
            out[2 * t] = c.real();
            out[2 * t + 1] = c.imag();
            finite = finite && std::isfinite(c.real()) && std::isfinite(c.imag());
        }
    } else {
        for (int r = 0; r < f.nread; r++)
            for (int t = 0; t < f.npoints; t++) {
                const cplx c = b->P_host[(size_t)(f.p_off + (int64_t)r * f.npoints + t)];
                out[(int64_t)r * cap + t] = c.real();
                finite = finite && std::isfinite(c.real());
            }
    }
    return finite ? EH_OK : fail(EH_ENAN, "non-finite readout (state diverged)");
}

int eh_batch_profile(eh_batch* b, int nsteps, double* out5) {
    return b ? batch_profile(b->b, nsteps, out5) : fail(EH_EINVAL, "NULL batch");
}

int eh_batch_run_eager(eh_batch* b) { return b ? batch_run_eager(b->b) : fail(EH_EINVAL, "NULL batch"); }

int eh_batch_autotune(eh_batch* b, int steps) {
    return b ? batch_autotune(b->b, steps) : fail(EH_EINVAL, "NULL batch");
}

int eh_batch_select(eh_batch* b, int group, int cand) {
    return b ? batch_select(b->b, group, cand) : fail(EH_EINVAL, "NULL batch");
}

int64_t eh_batch_tuning_report(const eh_batch* b, char* buf, int64_t cap) {
    if (!b) return fail(EH_EINVAL, "NULL batch");
    const std::string r = batch_tuning_report(b->b);
    if (buf && cap > 0) {
        const size_t n = std::min<size_t>(r.size(), (size_t)cap - 1);
        std::memcpy(buf, r.data(), n);
        buf[n] = 0;
    }
    return (int64_t)r.size();
}

int eh_batch_enable_adjoint(eh_batch* b, int checkpoint_every) {
    return b ? batch_enable_adjoint(b->b, checkpoint_every) : fail(EH_EINVAL, "NULL batch");
}

int eh_batch_backward(eh_batch* b, const double* dP, int64_t n_dP) {
    return (b && dP) ? batch_backward(b->b, dP, n_dP) : fail(EH_EINVAL, "NULL argument");
}

int eh_batch_set_chunk(eh_batch* b, int chunk) {
    return b ? batch_set_chunk(b->b, chunk) : fail(EH_EINVAL, "NULL batch");
}
int eh_batch_start(eh_batch* b) { return b ? batch_start(b->b) : fail(EH_EINVAL, "NULL batch"); }
int eh_batch_advance(eh_batch* b, int nsteps) {
    return b ? batch_advance(b->b, nsteps) : fail(EH_EINVAL, "NULL batch");
}

int eh_batch_plan(int n, const eh_instance* inst, int checkpoint_every, int64_t* o) {
    if (!o) return fail(EH_EINVAL, "eh_batch_plan: NULL argument");
    BatchPlan p;
    const int rc = plan_batch(n, inst, checkpoint_every, p);
    if (rc) return rc;
    o[0] = p.total_ados;
    o[1] = p.total_elems;
    o[2] = p.max_steps;
    o[3] = p.forward_bytes;
    o[4] = p.adjoint_replay_bytes;
    o[5] = p.adjoint_traj_bytes;
    return EH_OK;
}

int eh_batch_get_grad(eh_batch* b, int i, double* dH, double* dX4, double* dalpha, double* dalpha_t,
                      double* dgamma) {
    return b ? batch_get_grad(b->b, i, dH, dX4, dalpha, dalpha_t, dgamma) : fail(EH_EINVAL, "NULL batch");
}

int eh_batch_get_rho(eh_batch* b, int i, double* rho_out) {
    return b ? batch_get_rho(b->b, i, rho_out) : fail(EH_EINVAL, "NULL batch");
}

int eh_batch_dims(const eh_batch* bb, int64_t* o) {
    if (!bb || !bb->b || !o) return fail(EH_EINVAL, "NULL argument");
    const Batch* b = bb->b;
    o[0] = b->n_inst;
    o[1] = b->total_ados;
    o[2] = b->total_elems;
    o[3] = b->max_steps;
    o[4] = b->chunk;
    return EH_OK;
}

int64_t eh_batch_ado_updates(const eh_batch* bb) { return (bb && bb->b) ? bb->b->ado_updates : 0; }

static int solve_one(eh_ctx* ctx, const eh_instance& in, double* out, int64_t cap, int64_t* n_points,
                     double* rho_out) {
    eh_batch* b = nullptr;
    int rc = eh_batch_create(ctx, 1, &in, &b);
    if (rc) return rc;
    if ((rc = eh_batch_run(b)) == 0 && (rc = eh_batch_sync(b)) == 0) {
        rc = eh_batch_get_output(b, 0, out, cap, n_points);
        if ((rc == 0 || rc == EH_ENAN) && rho_out) {
            const int rc2 = eh_batch_get_rho(b, 0, rho_out);
            if (rc2) rc = rc2;
        }
    }
    eh_batch_destroy(b);
    return rc;
}

int eh_2des_solve(eh_ctx* ctx, const eh_system* sys, const double* H, const double* X4, int nv1, int nv2,
                  int nv3, double* P_out, int64_t cap, int64_t* n_points, double* rho_out) {
    eh_instance in;
    std::memset(&in, 0, sizeof(in));
    in.sys = sys;
    in.H = H;
    in.X4 = X4;
    in.nv[0] = nv1;
    in.nv[1] = nv2;
    in.nv[2] = nv3;
    in.mode = EH_MODE_2DES;
    return solve_one(ctx, in, P_out, cap, n_points, rho_out);
}

int eh_2des_solve_noalloc(eh_ctx* ctx, const eh_system* sys, const double* H, const double* X4, int nv1, int nv2,
                          int nv3, double* P_out, int64_t cap, int64_t* n_points, double* rho_out) {
    eh_instance in;
    std::memset(&in, 0, sizeof(in));
    in.sys = sys;
    in.H = H;
    in.X4 = X4;
    in.nv[0] = nv1;
    in.nv[1] = nv2;
    in.nv[2] = nv3;
    in.mode = EH_MODE_2DES;
    return solve_2des_noalloc(ctx, in, P_out, cap, n_points, rho_out);
}

int eh_dynamics_solve(eh_ctx* ctx, const eh_system* sys, const double* H, const int* sites, int n_sites,
                      double* pop_out, int64_t cap, int64_t* n_points, double* rho_out) {
    eh_instance in;
    std::memset(&in, 0, sizeof(in));
    in.sys = sys;
    in.H = H;
    in.mode = EH_MODE_DYNAMICS;
    in.sites = sites;
    in.n_sites = n_sites;
    return solve_one(ctx, in, pop_out, cap, n_points, rho_out);
}

int eh_2des_key_run(eh_ctx* ctx, const char* path, double* P_out, int64_t cap, int64_t* n_points) {
    eh_problem* pr = nullptr;
    int rc = eh_problem_from_key(path, EH_MODE_2DES, &pr);
    if (rc) return rc;
    std::unique_ptr<eh_problem, int (*)(eh_problem*)> guard(pr, eh_problem_destroy);
    const Problem& k = pr->p;
    const int ns = k.n_sample, N = k.sys_size;
    if (ns <= 0) return fail(EH_EPARSE, "DISORDER block missing or n_sample = 0 (the reference averages over no samples)");
    std::vector<cplx> H, X;
    if ((rc = draw_samples(k, ns, H, X))) return rc;
    static const int nvs[3][3] = {{1, 1, 1}, {1, 1, 0}, {1, 0, 1}};
    std::vector<eh_instance> inst((size_t)3 * ns);
    for (int s = 0; s < ns; s++)
        for (int v = 0; v < 3; v++) {
            eh_instance& q = inst[(size_t)3 * s + v];
            std::memset(&q, 0, sizeof(q));
            q.sys = &pr->sys;
            q.H = reinterpret_cast<const double*>(H.data() + (size_t)s * N * N);
            q.X4 = reinterpret_cast<const double*>(X.data() + (size_t)s * 4 * N * N);
            for (int c = 0; c < 3; c++) q.nv[c] = nvs[v][c];
            q.mode = EH_MODE_2DES;
        }
    eh_batch* b = nullptr;
    if ((rc = eh_batch_create(ctx, 3 * ns, inst.data(), &b))) return rc;
    std::unique_ptr<eh_batch, int (*)(eh_batch*)> bg(b, eh_batch_destroy);
    if ((rc = eh_batch_run(b)) || (rc = eh_batch_sync(b))) return rc;
    const int np = b->b->info[0].npoints;
    if (n_points) *n_points = np;
    if (!P_out) return EH_OK;
    if (np > cap) return fail(EH_ERANGE, "output capacity too small");
    std::vector<double> acc((size_t)2 * np, 0.0), p1((size_t)2 * np), p2((size_t)2 * np), p3((size_t)2 * np);
    int status = EH_OK;
    for (int s = 0; s < ns; s++) {
        int64_t n1 = 0;
        const int r1 = eh_batch_get_output(b, 3 * s, p1.data(), np, &n1);
        const int r2 = eh_batch_get_output(b, 3 * s + 1, p2.data(), np, &n1);
        const int r3 = eh_batch_get_output(b, 3 * s + 2, p3.data(), np, &n1);
        for (int r : {r1, r2, r3})
            if (r && r != EH_ENAN) return r;
            else if (r == EH_ENAN) status = EH_ENAN;
        // P_s = p1 - p2 - p3 (CPU_2DES.cpp:73-78); P += P_s / n_sample (:83-89)
        for (int t = 0; t < np; t++)
            for (int c = 0; c < 2; c++) {
                const double v = p1[(size_t)2 * t + c] - p2[(size_t)2 * t + c] - p3[(size_t)2 * t + c];
                acc[(size_t)2 * t + c] += v / ns;
            }
    }
    std::memcpy(P_out, acc.data(), acc.size() * sizeof(double));
    return status ? fail(status, "non-finite polarization") : EH_OK;
}

int eh_dynamics_key_run(eh_ctx* ctx, const char* path, const int* sites, int n_sites, double* pop_out,
                        int64_t cap, int64_t* n_points) {
    eh_problem* pr = nullptr;
    int rc = eh_problem_from_key(path, EH_MODE_DYNAMICS, &pr);
    if (rc) return rc;
    std::unique_ptr<eh_problem, int (*)(eh_problem*)> guard(pr, eh_problem_destroy);
    const Problem& k = pr->p;
    // population length as CPU_dynamics.cpp:26 computes it (float expression -> size)
    const float lenf = (k.t_end - k.t_start) / k.print_step + 1;
    const int64_t len = lenf > 0 ? (int64_t)lenf : 0;
    if (n_points) *n_points = len;
    if (!pop_out) return EH_OK;
    if (len > cap) return fail(EH_ERANGE, "output capacity too small");
    std::vector<cplx> H(k.Hal);
    int64_t np = 0;
    eh_system sys = pr->sys;
    // first pass for the number of readouts
    Schedule sch;
    const int nv0[3] = {0, 0, 0};
    if ((rc = build_schedule(sys, EH_MODE_DYNAMICS, nv0, sch))) return rc;
    std::vector<double> tmp((size_t)std::max(1, n_sites) * std::max(1, sch.npoints));
    rc = eh_dynamics_solve(ctx, &sys, reinterpret_cast<const double*>(H.data()), sites, n_sites, tmp.data(),
                           std::max(1, sch.npoints), &np, nullptr);
    if (rc && rc != EH_ENAN) return rc;
    for (int s = 0; s < n_sites; s++)
        for (int64_t t = 0; t < len; t++)
            pop_out[(int64_t)s * cap + t] = t < np ? tmp[(size_t)s * std::max(1, sch.npoints) + (size_t)t] : 0.0;
    return rc;
}

}  // extern "C"
