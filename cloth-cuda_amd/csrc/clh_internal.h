#pragma once
// Internal declarations shared by the translation units of libclothhip.
#include "clothhip.h"

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

namespace clh {

// Records `msg` as this thread's last error and returns `status`.
int fail(int status, const std::string& msg);
// Maps a HIP error to CLH_ERR_DEVICE with the call text; CLH_OK on success.
int hip_status(hipError_t e, const char* what);

#define CLH_HIP(call)                                                  \
    do {                                                               \
        const int _s = ::clh::hip_status((call), #call);               \
        if (_s != CLH_OK) return _s;                                   \
    } while (0)

#define CLH_TRY(call)                                                  \
    do {                                                               \
        const int _s = (call);                                         \
        if (_s != CLH_OK) return _s;                                   \
    } while (0)

// Device-side control of a PD frame run to a residual tolerance. Each
// iteration kernel reads the Chebyshev state of its iteration from
// slot[k & 1] and (block 0) writes the next iteration's into the other slot;
// the residual reduction after a check iteration records the history,
// restarts the Chebyshev sequence when the residual grew since the previous
// check, and raises `stop` on convergence, after which every later kernel of
// the frame returns at once. The host therefore needs no round trip per check.
struct CtlSlot {
    float omega;  // Chebyshev weight of this iteration
    int   blend;  // apply the blend
    int   ck;     // iterations since the (re)start of the sequence
    int   k;      // iteration index in the frame
};
struct FrameCtl {
    CtlSlot slot[2];
    float  rho;             // spectral radius estimate in use this frame
    int    use_cheb;
    int    delay;
    int    stop;            // converged or non-finite: later kernels are no-ops
    int    iters;           // iterations run when stop was raised
    int    checks;          // residual evaluations recorded
    int    restarts;        // Chebyshev restarts this frame
    int    converged;
    int    nonfinite;
    int    barrier_error;   // a grid barrier of the persistent kernel timed out
    double last_r2;         // residual^2 at the previous check (-1: none)
    double tol2;            // tolerance^2
    int    stall_checks;    // > 0: stop, not converged, after this many checks in a row
                            // without the relative residual dropping by 10 %
    int    since_drop;      // checks since the relative residual last dropped by 10 %
    double best_rel2;       // best (r / b)^2 of the frame so far
    double stall_ref;       // (r / b)^2 at the last 10 % drop
    int    stalled;         // the frame stopped on the stall rule
    int    restart_rule;    // CLH_RESTART_* (ctl_restart_due)
    float  restart_factor;  // CLH_RESTART_SEQ_MIN: growth factor over the sequence's minimum
    double seq_min_r2;      // smallest residual^2 of the current sequence (-1: none yet)
    // host-visible copies of this block (pinned, coherent): after check c
    // the reduction writes the block to mirror[c & 1], so the host reads the
    // control state of a chunk without a copy on the stream (null: no copies)
    FrameCtl* mirror;
};

// Per-cloth control of a batch run to a tolerance: each cloth has its own
// Chebyshev sequence, restarts, rho, convergence and stall test, so cloths do
// not steer one another. A cloth that stops (converged, stalled or non-finite)
// is frozen: its tiles keep x_{k+1} = x_k for the rest of the frame, and the
// batch frame ends when no cloth runs. Tiles read slot[par] of their cloth;
// the cloth's first tile writes slot[par ^ 1].
struct ClothCtl {
    CtlSlot slot[2];
    float  rho;
    int    state;       // CLH_CLOTH_*
    int    restarts;
    int    iters;       // iterations when the cloth stopped (0 while it runs)
    int    since_drop;
    int    pad;
    double last_r2;
    double best_rel2;
    double stall_ref;
    double r2, b2;      // residual^2 and ||b||^2 at the cloth's last check
    double seq_min_r2;  // smallest residual^2 of the cloth's current sequence (-1: none yet)
};

// Device buffer owned by the solver.
template <typename T>
struct DevBuf {
    T*     ptr   = nullptr;
    size_t count = 0;
    int resize(size_t n)
    {
        if (n <= count && ptr) return CLH_OK;
        release();
        if (n == 0) return CLH_OK;
        void* p = nullptr;
        if (hipMalloc(&p, n * sizeof(T)) != hipSuccess)
            return fail(CLH_ERR_ALLOC, "hipMalloc of solver scratch failed");
        ptr = static_cast<T*>(p);
        count = n;
        return CLH_OK;
    }
    int upload(const std::vector<T>& h)
    {
        CLH_TRY(resize(h.size()));
        if (!h.empty())
            CLH_HIP(hipMemcpy(ptr, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
        return CLH_OK;
    }
    void release()
    {
        if (ptr) (void)hipFree(ptr);
        ptr = nullptr;
        count = 0;
    }
};

}  // namespace clh
