// noalloc.hip -- one 2DES propagation that allocates no device memory.
//
// eh_2des_solve (twoD_spectrum_solver, etoheom-cpu/lib/dynamics.cpp:298-432) builds a
// batch of one instance: hipMalloc of its states, tables and readouts, a HIP graph.
// eh_2des_solve_noalloc runs the same propagation for callers that may not allocate
// device memory (serving loops, harnesses that forbid hipMalloc): the same host-side
// tables (build_host_batch) and the same stage kernels, launched one by one as the
// per-launch path does (prologue, four stages, counter tick per step; bit for bit the
// graph's results), with every device buffer carved out of `scratch`, one static
// device array of kScratchBytes per GPU.  The host copies its tables in and the
// readouts (and the final state) out with plain hipMemcpy.
//
// The array's host-side name is only a stand-in: its device address on each GPU comes
// from hipGetSymbolAddress, called once per device at context creation (scratch_setup,
// from ctx_init) with that device current, and only that address is used.  Every call
// on one GPU shares the array and leaves its contents behind; reset_state clears the
// state and the column copies at the start of each call, and every other buffer is
// written before it is read.  Calls that overlapped on one GPU would race on it, so a
// per-device lock is held for a whole call: concurrent calls from threads of this
// process run one after another.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "engine_impl.hpp"

namespace eh {

constexpr size_t kScratchBytes = size_t(EH_SCRATCH_BYTES);
constexpr int kMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {
std::mutex g_dev_mu[kMaxDevices];
unsigned char* g_scratch[kMaxDevices];   // device address of `scratch` on each device

// bump allocator over the scratch array: 256-byte aligned offsets
struct Carve {
    size_t off = 0;
    template <typename T>
    size_t take(size_t count) {
        const size_t o = off;
        off += (count * sizeof(T) + 255) & ~size_t(255);
        return o;
    }
};
}  // namespace

// Device address of `scratch` on `device` (which must be current), once per device.
int scratch_setup(int device) {
    if (device < 0 || device >= kMaxDevices) return fail(EH_EINVAL, "scratch_setup: device index out of range");
    std::lock_guard<std::mutex> lk(g_dev_mu[device]);
    if (g_scratch[device]) return EH_OK;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        return fail(EH_EHIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch[device] = static_cast<unsigned char*>(p);
    return EH_OK;
}

int solve_2des_noalloc(eh_ctx* ctx, const eh_instance& in, double* P_out, int64_t cap, int64_t* n_points,
                       double* rho_out) {
    if (!ctx) return fail(EH_EINVAL, "eh_2des_solve_noalloc: NULL context");
    const int dev = ctx->device;
    if (dev < 0 || dev >= kMaxDevices) return fail(EH_EINVAL, "eh_2des_solve_noalloc: device index out of range");
    Batch b;   // sizes and per-instance info only: no device state
    HostBatch hb;
    int rc = build_host_batch(ctx, 1, &in, &b, hb);
    if (rc) return rc;
    const Batch::Info& f = b.info[0];
    if (n_points) *n_points = f.npoints;
    if (P_out && f.npoints > cap) return fail(EH_ERANGE, "eh_2des_solve_noalloc: output capacity too small");

    // the buffers of batch_create, one copy of V (no fused steps), carved in order
    Carve cv;
    const size_t o_inst = cv.take<InstDev>(hb.inst.size()), o_ti = cv.take<int>(hb.t_inst.size()),
                 o_ta = cv.take<int>(hb.t_ado0.size()), o_tn = cv.take<int>(hb.t_nado.size()),
                 o_up = cv.take<int>(hb.up.size()), o_dn = cv.take<int>(hb.dn.size()),
                 o_cx = cv.take<double>(hb.cx.size()), o_cy = cv.take<double>(hb.cy.size()),
                 o_pidx = cv.take<int>(hb.pidx.size()), o_vreal = cv.take<uint8_t>(hb.vreal.size()),
                 o_H = cv.take<double2>(hb.H.size()), o_Xp = cv.take<double2>(hb.Xp.size()),
                 o_field = cv.take<double2>(hb.field.size()), o_R = cv.take<double2>(hb.R.size()),
                 o_al = cv.take<double2>(hb.al.size()), o_alt = cv.take<double2>(hb.alt.size()),
                 o_damp = cv.take<double2>(hb.damp.size()), o_ctr = cv.take<int>(1),
                 o_V = cv.take<double2>(hb.H.size()), o_Vpad = cv.take<double>((size_t)hb.vp_elems),
                 o_P = cv.take<double2>((size_t)hb.p_elems), o_spc = cv.take<int>(hb.spcnt.size()),
                 o_spi = cv.take<short>(hb.spidx.size());
    size_t o_st[4], o_col[3];
    for (size_t& o : o_st) o = cv.take<double2>((size_t)hb.state_elems);
    for (size_t& o : o_col) o = cv.take<double2>((size_t)hb.col_elems);
    if (cv.off > kScratchBytes)
        return fail(EH_ERANGE, "eh_2des_solve_noalloc: the propagation needs " + std::to_string(cv.off) +
                                   " bytes of device scratch (" + std::to_string(f.M) + " ADOs of " +
                                   std::to_string(f.N) + "x" + std::to_string(f.N) + "), more than its fixed " +
                                   std::to_string(kScratchBytes) + "; use eh_2des_solve");

    HIPCHK(hipSetDevice(dev));
    if ((rc = scratch_setup(dev))) return rc;
    std::lock_guard<std::mutex> lk(g_dev_mu[dev]);
    unsigned char* const base = g_scratch[dev];
    auto put = [&](size_t off, const void* src, size_t bytes) -> int {
        if (bytes) HIPCHK(hipMemcpy(base + off, src, bytes, hipMemcpyHostToDevice));
        return EH_OK;
    };
#define EH_PUT(off, vec)                                                                       \
    if ((rc = put(off, (vec).data(), (vec).size() * sizeof((vec)[0])))) return rc;
    EH_PUT(o_inst, hb.inst) EH_PUT(o_ti, hb.t_inst) EH_PUT(o_ta, hb.t_ado0) EH_PUT(o_tn, hb.t_nado)
    EH_PUT(o_up, hb.up) EH_PUT(o_dn, hb.dn) EH_PUT(o_cx, hb.cx) EH_PUT(o_cy, hb.cy) EH_PUT(o_pidx, hb.pidx)
    EH_PUT(o_vreal, hb.vreal) EH_PUT(o_H, hb.H) EH_PUT(o_Xp, hb.Xp) EH_PUT(o_field, hb.field) EH_PUT(o_R, hb.R)
    EH_PUT(o_al, hb.al) EH_PUT(o_alt, hb.alt) EH_PUT(o_damp, hb.damp) EH_PUT(o_spc, hb.spcnt)
    EH_PUT(o_spi, hb.spidx)
#undef EH_PUT
    auto at = [&](size_t off) { return static_cast<void*>(base + off); };
    const InstDev* inst = static_cast<const InstDev*>(at(o_inst));
    int* ctr = static_cast<int*>(at(o_ctr));
    double2* st[4];
    double2* col[3];
    for (int q = 0; q < 4; q++) st[q] = static_cast<double2*>(at(o_st[q]));
    for (int q = 0; q < 3; q++) col[q] = static_cast<double2*>(at(o_col[q]));
    double2 *Y = st[0], *ACC = st[1], *S1 = st[2], *S2 = st[3];
    double2 *Yc = col[0], *S1c = col[1], *S2c = col[2];
    double2* P = static_cast<double2*>(at(o_P));

    StageArgs sa;
    std::memset(&sa, 0, sizeof(sa));
    sa.inst = inst;
    sa.tile_inst = static_cast<const int*>(at(o_ti));
    sa.tile_ado0 = static_cast<const int*>(at(o_ta));
    sa.tile_nado = static_cast<const int*>(at(o_tn));
    sa.step_ctr = ctr;
    sa.vreal = static_cast<const uint8_t*>(at(o_vreal));
    sa.V = static_cast<const double2*>(at(o_V));
    sa.Vpad = hb.vp_elems ? static_cast<const double*>(at(o_Vpad)) : nullptr;
    sa.nb_up = static_cast<const int*>(at(o_up));
    sa.nb_dn = static_cast<const int*>(at(o_dn));
    sa.cx = static_cast<const double*>(at(o_cx));
    sa.cy = static_cast<const double*>(at(o_cy));
    sa.alpha = static_cast<const double2*>(at(o_al));
    sa.alpha_t = static_cast<const double2*>(at(o_alt));
    sa.damp = static_cast<const double2*>(at(o_damp));
    sa.sp_cnt = static_cast<const int*>(at(o_spc));
    sa.sp_idx = static_cast<const short*>(at(o_spi));
    sa.gsign = 1.0;
    sa.y = Y;
    sa.acc = ACC;
    PrologueArgs pa;
    std::memset(&pa, 0, sizeof(pa));
    pa.inst = inst;
    pa.step_ctr = ctr;
    pa.H = static_cast<const double2*>(at(o_H));
    pa.Xp = static_cast<const double2*>(at(o_Xp));
    pa.field = static_cast<const double2*>(at(o_field));
    pa.pidx = static_cast<const int*>(at(o_pidx));
    pa.R = static_cast<const double2*>(at(o_R));
    pa.y = Y;
    pa.V = static_cast<double2*>(at(o_V));
    pa.Vpad = hb.vp_elems ? static_cast<double*>(at(o_Vpad)) : nullptr;
    pa.P = P;
    pa.readout = 1;

    // the RK4 stage chain of batch_run_eager: stage s reads in[s] and writes out[s]
    double2* const rin[4] = {Y, S1, S2, S1};
    double2* const rout[4] = {S1, S2, S1, Y};
    double2* const cin[4] = {Yc, S1c, S2c, S1c};
    double2* const cout[4] = {S1c, S2c, S1c, Yc};
    const hipStream_t s = nullptr;   // the null stream: no stream is created
    launch_reset_raw(s, Y, hb.state_elems, Yc, hb.col_elems, ctr, inst, 1);
    const Group& g = hb.groups[0];
    for (int k = 0; k < f.nsteps; k++) {
        hipLaunchKernelGGL(prologue, dim3(1), dim3(256), 0, s, pa);
        for (int q = 0; q < 4; q++) {
            StageArgs a = sa;
            a.stage = q;
            a.rin = rin[q];
            a.out = rout[q];
            a.cin = cin[q];
            a.cout = cout[q];
            if ((rc = launch_stage(g, a, s))) return rc;
        }
        hipLaunchKernelGGL(tick, dim3(1), dim3(1), 0, s, ctr, 1);
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(s));
    bool finite = true;
    if (P_out && f.npoints) {
        HIPCHK(hipMemcpy(P_out, P, (size_t)f.npoints * sizeof(double2), hipMemcpyDeviceToHost));
        for (int64_t t = 0; t < 2 * (int64_t)f.npoints; t++) finite = finite && std::isfinite(P_out[t]);
    }
    if (rho_out)
        HIPCHK(hipMemcpy(rho_out, Y, (size_t)f.M * f.N * f.N * sizeof(double2), hipMemcpyDeviceToHost));
    return finite ? EH_OK : fail(EH_ENAN, "non-finite readout (state diverged)");
}

}  // namespace eh
