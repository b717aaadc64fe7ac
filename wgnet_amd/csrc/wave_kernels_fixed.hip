// Resident adjoint kernels specialised for the reference grid (70 x 70
// model, 20-cell C-PML: 114 x 114 padded), one workgroup per unit, and the
// adjoint time-split work queue; the device templates are in
// fixed_kernels.h, the forward kernels in wave_kernels_fixed_fwd.hip.
#include "queue_fx.h"

namespace wg {

namespace fx {
template <class G, bool DWAV>
__global__ __launch_bounds__(NT) void k_adjoint_fx(AdjParams P) {
    extern __shared__ float lds[];
    adjoint_run<G, DWAV>(P, lds, blockIdx.x, P.g.nt, 0, nullptr, nullptr);
}

template <class G>
__global__ __launch_bounds__(NT) void k_queue_adj_fx(QueueParams Q) {
    extern __shared__ float lds[];
    __shared__ int task_s;
    const int parts = Q.parts, units = Q.units, nt = Q.a.g.nt;
    int j, u;
    while (next_task(Q.next_a, Q.done_a, parts * units, units, Q.err, Q.spin_limit, &task_s, j, u)) {
        float* st = Q.astate + (size_t)u * AState<G>::FLOATS;
        const int sg = parts - 1 - j;
        const int t_lo = (int)((long long)nt * sg / parts), t_hi = (int)((long long)nt * (sg + 1) / parts);
        adjoint_run<G, false>(Q.a, lds, u, t_hi, t_lo, j > 0 ? st : nullptr, j < parts - 1 ? st : nullptr);
        task_done(Q.done_a, units, j, u);
    }
}

}  // namespace fx

// ---------------------------------------------------------------------------
// Dispatch: the compiled geometry, or false when the grid is another one.
// ---------------------------------------------------------------------------

size_t adjoint_lds_bytes_fx(const Grid& g, const Geometry& geo) {
    if (!is_g70(g)) return 0;
    return 4ull * (fx::G70::A_TAIL + g.nrec + 8 + geo.band_rows * fx::G70::NX + 1 + 2 * geo.n_ent);
}

bool launch_adjoint_fx(const AdjParams& p, int units, hipStream_t st, hipError_t* err) {
    if (!is_g70(p.g)) return false;
    const size_t lds = adjoint_lds_bytes_fx(p.g, p.geo);
    void (*fn)(AdjParams) = p.dwav ? fx::k_adjoint_fx<fx::G70, true> : fx::k_adjoint_fx<fx::G70, false>;
    hipError_t e = hipFuncSetAttribute((const void*)fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(fn, dim3(units), dim3(fx::NT), lds, st, p);
        e = hipGetLastError();
    }
    *err = e;
    return true;
}

bool queue_state_fx(const Grid& g, size_t* fwd_floats, size_t* adj_floats) {
    if (!is_g70(g)) return false;
    *fwd_floats = fx::FState<fx::G70>::FLOATS;
    *adj_floats = fx::AState<fx::G70>::FLOATS;
    return true;
}

bool launch_queue_fx(const QueueParams& q, bool adjoint, int workgroups, hipStream_t st, hipError_t* err) {
    if (!is_g70(q.f.g)) return false;
    if (!adjoint) return launch_queue_fwd_fx(q, workgroups, st, err);
    void (*fn)(QueueParams) = fx::k_queue_adj_fx<fx::G70>;
    const size_t lds = adjoint_lds_bytes_fx(q.a.g, q.a.geo);
    hipError_t e = hipFuncSetAttribute((const void*)fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e == hipSuccess && workgroups > 0) {
        hipLaunchKernelGGL(fn, dim3(workgroups), dim3(fx::NT), lds, st, q);
        e = hipGetLastError();
    }
    *err = e;
    return true;
}

}  // namespace wg
