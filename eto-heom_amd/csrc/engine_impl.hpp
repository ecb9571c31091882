// engine_impl.hpp -- host-side internals of the batch engine shared by engine.hip
// and adjoint.hip (device buffers, kernel groups, the per-batch state).
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "device.hpp"
#include "engine.hpp"

namespace eh {

#define HIPCHK(expr)                                                                   \
    do {                                                                               \
        hipError_t e_ = (expr);                                                        \
        if (e_ != hipSuccess)                                                          \
            return fail(EH_EHIP, std::string(#expr) + ": " + hipGetErrorString(e_));   \
    } while (0)

template <typename T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
    int alloc(size_t count) {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = count;
        if (count == 0) count = 1;
        hipError_t e = hipMalloc((void**)&p, count * sizeof(T));
        if (e != hipSuccess) {
            p = nullptr;
            return fail(e == hipErrorOutOfMemory ? EH_ENOMEM : EH_EHIP,
                        std::string("hipMalloc: ") + hipGetErrorString(e));
        }
        return EH_OK;
    }
    int upload(const std::vector<T>& h) {
        int rc = alloc(h.size());
        if (rc) return rc;
        if (!h.empty()) HIPCHK(hipMemcpy(p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
        return EH_OK;
    }
};

// One way of launching a group's stage: kernel variant and tiling.
struct LaunchCfg {
    int nt;                   // 1..16: stage_small<nt>; -32 / -64: stage_mfma; -65: stage_mfma64;
                              // -33: stage_mfma_pipe; -128: stage_mfma_wide; -67: stage_sparse;
                              // -69: stage_sparse_rows (stage 1: stage_sparse unless rows1);
                              // 0: stage_generic
    int ge;                   // stage_small: state elements per tile (256, 512 or 1024);
                              // stage_mfma_wide: workgroups per ADO
    bool cols;                // keep coupled-column copies (contiguous column gathers)
    bool stream;              // non-temporal state stores (StageArgs::stream)
    bool late = false;        // stage_small<N, 512, true>: state operand loaded after the barrier
    bool rows1 = false;       // stage_sparse_rows: stage 1 too (its exchange fits the LDS past rho)
    int tile_begin, tile_count;
    size_t shmem;
};

// The instances of one N class (N <= 16 each its own class; 16 < N <= 32; 32 < N <= 64;
// 64 < N <= 128).
struct Group {
    int nt;                   // base configuration (cand[0]); the adjoint kernels use its tiling
    int tile_begin, tile_count;
    size_t shmem;
    int maxnn, maxkn, maxd;   // largest N^2, 2 K N and d of the group's instances
    int maxsp;                // largest row count of the instances' sparsity patterns of V
    int persist;              // grid of the persistent kernels (workgroups)
    double h, h2, hb[4];      // RK coefficients of the pipelined variant (one time step)
    int inst_begin, inst_count;   // the group's instances, a range of BatchImpl::ginst
    std::vector<LaunchCfg> cand;  // forward launch candidates, cand[0] = the base configuration
    std::vector<double> cand_ms;  // autotuner timings per candidate (ms per RK4 step), empty if not tuned
    int sel = 0;                  // the candidate the forward graph launches
    bool adj_pair = false;        // adj_mfma64 runs two ADOs per workgroup (set by the adjoint's enable)
    bool adj_global = false;      // adj_generic reads lambda and Y in place (they do not fit the LDS)
};

// Host side of a batch (build_host_batch): the tables batch_create uploads, the
// groups with their launch candidates, and the element counts of the device buffers.
struct HostBatch {
    std::vector<InstDev> inst;
    std::vector<int> t_inst, t_ado0, t_nado, up, dn, pidx, ginst;
    std::vector<uint8_t> ado, vreal;
    std::vector<double> cx, cy;
    std::vector<double2> H, Xp, field, R, al, alt, damp;
    std::vector<long long> ado0;
    std::vector<Group> groups;
    int64_t state_elems = 0;   // complex elements of each state buffer (Y, ACC, S1, S2)
    int64_t p_elems = 0;       // complex readouts
    int64_t col_elems = 0;     // complex elements of each coupled-column copy
    int64_t vp_elems = 0;      // doubles of one copy of the padded V planes
    std::vector<int> spcnt;    // sparsity pattern of V (rows of distinct patterns, InstDev::sp_off)
    std::vector<short> spidx;
};
int build_host_batch(eh_ctx* ctx, int n, const eh_instance* in, Batch* b, HostBatch& hb);
// reset_state + set_ground of engine.hip on raw buffers: every state element and
// column copy to zero, rho_0 = |0><0| per instance, step counter to 0
void launch_reset_raw(hipStream_t s, double2* y, long long total, double2* yc, long long ctotal, int* ctr,
                      const InstDev* inst, int n_inst);

// noalloc.hip: the device address of its static scratch array on `device` (current),
// looked up once per device (ctx_init)
int scratch_setup(int device);

struct AdjState;   // adjoint.hip
void adj_destroy(AdjState* a);

struct BatchImpl {
    hipStream_t stream = nullptr;
    hipGraphExec_t exec = nullptr;
    hipGraph_t graph = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    DevBuf<InstDev> inst;
    DevBuf<int> tile_inst, tile_ado0, tile_nado, step_ctr, nb_up, nb_dn, pidx;
    DevBuf<uint8_t> vreal;
    DevBuf<double> cx, cy;
    DevBuf<double2> V, H, Xp, field, R, P, alpha, alpha_t, damp, Y, ACC, S1, S2;
    DevBuf<FusedDev> fz;          // the fused steps' pointers (V holds two copies, by step parity)
    DevBuf<double2> Yc, S1c, S2c;        // coupled-column copies of Y, S1, S2
    DevBuf<double> Vpad;                 // padded V planes of the N > 32 instances
    DevBuf<long long> ado0_off;
    DevBuf<int> sp_cnt;                  // sparsity pattern of V per instance row (StageArgs::sp_cnt)
    DevBuf<short> sp_idx;
    DevBuf<int> ginst;                   // instance indices, grouped
    std::vector<Group> groups;
    std::vector<hipStream_t> side;       // one capture branch per group beyond the first
    hipEvent_t fork_ev = nullptr;
    std::vector<hipEvent_t> join_ev;
    int nchunks = 0;
    bool ran = false;         // a run (or an advance) has been issued since creation
    bool started = false;     // eh_batch_start: states reset, steps run by eh_batch_advance
    int pos = 0;              // steps run since the last start
    // host copies the adjoint needs to build its tables
    std::vector<InstDev> hinst;
    std::vector<int> hup, hdn;
    std::vector<double> hcx, hcy;
    std::vector<double2> hal, halt, hdamp;
    std::vector<uint8_t> hado;           // multi-indices, same (ADO, p) layout as the tables
    std::vector<int> htile_inst;         // instance of every tile
    AdjState* adj = nullptr;

    ~BatchImpl() {
        if (adj) adj_destroy(adj);
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        if (fork_ev) (void)hipEventDestroy(fork_ev);
        for (hipEvent_t e : join_ev) (void)hipEventDestroy(e);
        for (hipStream_t q : side) (void)hipStreamDestroy(q);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

int launch_cfg(const Group& g, const LaunchCfg& c, const StageArgs& base, hipStream_t s);
// the group's selected candidate
inline int launch_stage(const Group& g, const StageArgs& base, hipStream_t s) {
    return launch_cfg(g, g.cand[(size_t)g.sel], base, s);
}
// stage st of step k: input and output buffers of the RK4 scheme
StageArgs stage_io(const StageArgs& sa, const BatchImpl* im, int k, int st);
int adj_forward(Batch* b);
StageArgs stage_args(const BatchImpl* im);
PrologueArgs prologue_args(const BatchImpl* im);

}  // namespace eh
