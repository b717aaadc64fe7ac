// Resident forward kernels specialised for the reference grid (70 x 70
// model, 20-cell C-PML: 114 x 114 padded), one workgroup per unit, and the
// forward time-split work queue (queue_fx.h); the device templates are in
// fixed_kernels.h.  A translation unit of its own so that the forward bodies
// are scheduled with the strategy that suits them (the build passes it).
#include "queue_fx.h"

namespace wg {

namespace fx {
template <class G, bool STORE, bool BORN>
__global__ __launch_bounds__(NT) void k_forward_fx(FwdParams P) {
    extern __shared__ float lds[];
    forward_run<G, STORE, BORN>(P, lds, blockIdx.x, 0, P.g.nt, nullptr, nullptr);
}

template <class G>
__global__ __launch_bounds__(NT) void k_queue_fwd_fx(QueueParams Q) {
    extern __shared__ float lds[];
    __shared__ int task_s;
    const int parts = Q.parts, units = Q.units, nt = Q.f.g.nt;
    int p, u;
    while (next_task(Q.next_f, Q.done_f, parts * units, units, Q.err, Q.spin_limit, &task_s, p, u)) {
        float* st = Q.fstate + (size_t)u * FState<G>::FLOATS;
        const int t0 = (int)((long long)nt * p / parts), t1 = (int)((long long)nt * (p + 1) / parts);
        forward_run<G, true, false>(Q.f, lds, u, t0, t1, p > 0 ? st : nullptr, p < parts - 1 ? st : nullptr);
        task_done(Q.done_f, units, p, u);
    }
}

}  // namespace fx

size_t forward_lds_bytes_fx(const Grid& g) { return is_g70(g) ? 4ull * fx::G70::F_TOTAL : 0; }

bool launch_forward_fx(const FwdParams& p, int units, bool store, bool born, hipStream_t st, hipError_t* err) {
    if (!is_g70(p.g)) return false;
    const size_t lds = forward_lds_bytes_fx(p.g);
    void (*fn)(FwdParams) = born ? fx::k_forward_fx<fx::G70, false, true>
                                 : (store ? fx::k_forward_fx<fx::G70, true, false> : fx::k_forward_fx<fx::G70, false, false>);
    hipError_t e = hipFuncSetAttribute((const void*)fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(fn, dim3(units), dim3(fx::NT), lds, st, p);
        e = hipGetLastError();
    }
    *err = e;
    return true;
}

bool launch_queue_fwd_fx(const QueueParams& q, int workgroups, hipStream_t st, hipError_t* err) {
    if (!is_g70(q.f.g)) return false;
    void (*fn)(QueueParams) = fx::k_queue_fwd_fx<fx::G70>;
    const size_t lds = forward_lds_bytes_fx(q.f.g);
    hipError_t e = hipFuncSetAttribute((const void*)fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e == hipSuccess && workgroups > 0) {
        hipLaunchKernelGGL(fn, dim3(workgroups), dim3(fx::NT), lds, st, q);
        e = hipGetLastError();
    }
    *err = e;
    return true;
}

}  // namespace wg
