// engine.hpp -- host-side interface of the HIP batch engine (engine.hip).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "host.hpp"

struct eh_ctx {
    int device = 0;
    char name[256] = {0};
    int n_cu = 0;
    int lds_per_block = 0;
};

namespace eh {

struct BatchImpl;

// Everything one batched run needs, built once and replayed.
struct Batch {
    eh_ctx* ctx = nullptr;
    BatchImpl* impl = nullptr;
    int n_inst = 0;
    struct Info {
        int N = 0, M = 0, nsteps = 0, npoints = 0, nread = 0, mode = 0;
        int64_t state_off = 0, p_off = 0;
    };
    std::vector<Info> info;
    int64_t total_ados = 0, total_elems = 0;
    int max_steps = 0, chunk = 0;
    int64_t ado_updates = 0;
    std::vector<cplx> P_host;   // readout copy after sync
    bool synced = false;
};

int batch_create(eh_ctx* ctx, int n, const eh_instance* inst, Batch** out);
void batch_destroy(Batch* b);
int batch_run(Batch* b);
int batch_sync(Batch* b);
int batch_elapsed_ms(Batch* b, double* ms);
int batch_get_rho(Batch* b, int i, double* rho_out);
int batch_profile(Batch* b, int nsteps, double* out5);
int batch_run_eager(Batch* b);
int batch_set_chunk(Batch* b, int chunk);
int batch_start(Batch* b);
int batch_advance(Batch* b, int nsteps);
int batch_autotune(Batch* b, int steps);
int batch_select(Batch* b, int group, int cand);
std::string batch_tuning_report(const Batch* b);
int batch_enable_adjoint(Batch* b, int checkpoint_every);
int batch_backward(Batch* b, const double* dP, int64_t n_dP);
int batch_get_grad(Batch* b, int i, double* dH, double* dX4, double* dalpha, double* dalpha_t, double* dgamma);
int ctx_init(int device, eh_ctx* ctx);
// one 2DES propagation over the static device scratch array (noalloc.hip)
int solve_2des_noalloc(eh_ctx* ctx, const eh_instance& in, double* P_out, int64_t cap, int64_t* n_points,
                       double* rho_out);
int ctx_mem_info(eh_ctx* ctx, int64_t* free_bytes, int64_t* total_bytes);

}  // namespace eh
