// hc_batch.hip -- deferred single-step update for host-resident patches.
//
// The drop-in path for an unmodified ForestClaw host: patch_vt->
// single_step_update is called once per patch of a level with
// fclaw2d_single_step_buffer_data_t {total_count, iter}
// (src/fclaw2d_update_single_step.c:45-103, src/fclaw2d_update_single_step.h:47-55).
// Each call stages its patch into pinned host memory; when the buffer is full
// or the last patch of the level arrives, the staged tiles are copied to the
// device in one transfer, stepped by one batched launch (in place) and the
// updated interior cells are copied back to the callers' arrays; their ghost
// cells are left as given (the reference's step also rewrites them, but the
// next ghost fill overwrites them in any case).  The call that flushes returns the batch Courant number, every
// other call 0, which the caller's fmax reduction (:68) absorbs -- the same
// protocol as the reference's GPU solver (src/solvers/fc2d_cudaclaw/fc2d_cudaclaw.cpp:309-354).
#include <cstring>
#include <new>
#include <vector>

#include "hc_internal.hpp"

using hc::check_hip;
using hc::set_error;
using hc::is_claw5;
using hc::patch_from_host;
using hc::validate_physics;

namespace {

struct Staged {
    double* q;     // caller's patch array, T*meqn doubles
};

}  // namespace

struct hipclaw_batch {
    hipclaw_physics phys{};
    int cap = 0;
    hipStream_t stream = nullptr;
    long long T = 0;          // tile cells
    int esz = 8;
    void* h_q = nullptr;      // pinned staging, cap tiles (precision of the solver)
    void* h_aux = nullptr;
    double* h_dxdy = nullptr; // pinned
    double* h_cfl = nullptr;  // pinned, one value
    void* d_q = nullptr;
    void* d_aux = nullptr;
    double* d_dxdy = nullptr;
    double* d_cfl = nullptr;  // cap + 1 (last = batch max)
    std::vector<Staged> staged;
    double t = 0, dt = 0;
    bool have_dt = false;
    int src2 = HIPCLAW_SRC2_NONE;
};

namespace {

int flush(hipclaw_batch* b, double* maxcfl)
{
    const int n = (int)b->staged.size();
    *maxcfl = 0.0;
    if (n == 0) return HIPCLAW_OK;
    const int meqn = b->phys.meqn, maux = b->phys.maux;
    const size_t qb = (size_t)n * b->T * meqn * b->esz;
    int rc = check_hip(hipMemcpyAsync(b->d_q, b->h_q, qb, hipMemcpyHostToDevice, b->stream), "batch q H2D");
    if (!rc && maux > 0)
        rc = check_hip(hipMemcpyAsync(b->d_aux, b->h_aux, (size_t)n * b->T * maux * b->esz,
                                      hipMemcpyHostToDevice, b->stream), "batch aux H2D");
    if (!rc)
        rc = check_hip(hipMemcpyAsync(b->d_dxdy, b->h_dxdy, sizeof(double) * 2 * n, hipMemcpyHostToDevice,
                                      b->stream), "batch dxdy H2D");
    if (!rc)
        rc = hipclaw_step2_batch(&b->phys, n, b->d_q, b->d_q, maux > 0 ? b->d_aux : nullptr, b->d_dxdy,
                                 nullptr, n, b->dt, b->d_cfl, nullptr, b->stream);
    if (!rc && b->src2)
        rc = hipclaw_src2_batch(&b->phys, b->src2, n, b->d_q, b->d_aux, nullptr, nullptr, b->t, b->dt, b->stream);
    if (!rc) rc = hipclaw_max_reduce(b->d_cfl, n, b->d_cfl + b->cap, b->stream);
    if (!rc) rc = check_hip(hipMemcpyAsync(b->h_q, b->d_q, qb, hipMemcpyDeviceToHost, b->stream), "batch q D2H");
    if (!rc)
        rc = check_hip(hipMemcpyAsync(b->h_cfl, b->d_cfl + b->cap, sizeof(double), hipMemcpyDeviceToHost,
                                      b->stream), "batch cfl D2H");
    if (!rc) rc = check_hip(hipStreamSynchronize(b->stream), "batch sync");
    if (rc) {
        b->staged.clear();
        return rc;
    }
    // interior cells back to the callers' arrays; their ghost cells stay as given
    const long long per = b->T * meqn;
    const int NX = b->phys.mx + 4, NY = b->phys.my + 4;
    if (is_claw5(b->phys.solver)) {
        for (int k = 0; k < n; k++) {
            double* dst = b->staged[k].q;
            for (int j = 2; j < NY - 2; j++)
                for (int i = 2; i < NX - 2; i++)
                    for (int m = 0; m < meqn; m++) {
                        const long long o = ((long long)m * NY + j) * NX + i;
                        dst[((long long)j * NX + i) * meqn + m] =
                            b->esz == 8 ? static_cast<double*>(b->h_q)[k * per + o]
                                        : (double)static_cast<float*>(b->h_q)[k * per + o];
                    }
        }
        *maxcfl = *b->h_cfl;
        b->staged.clear();
        b->have_dt = false;
        return HIPCLAW_OK;
    }
    for (int k = 0; k < n; k++) {
        double* dst = b->staged[k].q;
        for (int m = 0; m < meqn; m++)
            for (int j = 2; j < NY - 2; j++) {
                const long long o = ((long long)m * NY + j) * NX + 2;
                if (b->esz == 8) {
                    std::memcpy(dst + o, static_cast<double*>(b->h_q) + k * per + o, sizeof(double) * (NX - 4));
                } else {
                    const float* src = static_cast<float*>(b->h_q) + k * per + o;
                    for (int i = 0; i < NX - 4; i++) dst[o + i] = (double)src[i];
                }
            }
    }
    *maxcfl = *b->h_cfl;
    b->staged.clear();
    b->have_dt = false;
    return HIPCLAW_OK;
}

}  // namespace

extern "C" {

int hipclaw_batch_create(const hipclaw_physics* phys, int capacity, void* stream, hipclaw_batch** out)
{
    int rc = validate_physics(phys);
    if (rc) return rc;
    if (capacity < 1 || !out) return set_error(HIPCLAW_ERR_ARG, "capacity < 1 or out is NULL");
    auto* b = new (std::nothrow) hipclaw_batch;
    if (!b) return set_error(HIPCLAW_ERR_NOMEM, "host allocation failed");
    b->phys = *phys;
    b->cap = capacity;
    b->stream = static_cast<hipStream_t>(stream);
    b->T = (long long)(phys->mx + 4) * (phys->my + 4);
    b->esz = phys->precision == HIPCLAW_FP64 ? 8 : 4;
    const size_t qb = (size_t)capacity * b->T * phys->meqn * b->esz;
    const size_t ab = (size_t)capacity * b->T * (phys->maux > 0 ? phys->maux : 1) * b->esz;
    rc = check_hip(hipHostMalloc(&b->h_q, qb, hipHostMallocDefault), "pinned q");
    if (!rc) rc = check_hip(hipHostMalloc(&b->h_aux, ab, hipHostMallocDefault), "pinned aux");
    if (!rc) rc = check_hip(hipHostMalloc((void**)&b->h_dxdy, sizeof(double) * 2 * capacity, hipHostMallocDefault), "pinned dxdy");
    if (!rc) rc = check_hip(hipHostMalloc((void**)&b->h_cfl, sizeof(double), hipHostMallocDefault), "pinned cfl");
    if (!rc) rc = check_hip(hipMalloc(&b->d_q, qb), "device q");
    if (!rc) rc = check_hip(hipMalloc(&b->d_aux, ab), "device aux");
    if (!rc) rc = check_hip(hipMalloc((void**)&b->d_dxdy, sizeof(double) * 2 * capacity), "device dxdy");
    if (!rc) rc = check_hip(hipMalloc((void**)&b->d_cfl, sizeof(double) * (capacity + 1)), "device cfl");
    if (rc) {
        hipclaw_batch_destroy(b);
        return rc;
    }
    b->staged.reserve(capacity);
    *out = b;
    return HIPCLAW_OK;
}

int hipclaw_batch_destroy(hipclaw_batch* b)
{
    if (!b) return HIPCLAW_OK;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    void* hp[] = {b->h_q, b->h_aux, b->h_dxdy, b->h_cfl};
    for (void* p : hp)
        if (p) (void)hipHostFree(p);
    void* dp[] = {b->d_q, b->d_aux, b->d_dxdy, b->d_cfl};
    for (void* p : dp)
        if (p) (void)hipFree(p);
    delete b;
    return HIPCLAW_OK;
}

int hipclaw_batch_set_src2(hipclaw_batch* b, int kind)
{
    if (!b) return set_error(HIPCLAW_ERR_ARG, "null batch");
    if (kind != HIPCLAW_SRC2_NONE && kind != HIPCLAW_SRC2_AXISYMMETRIC)
        return set_error(HIPCLAW_ERR_ARG, "unknown source term");
    if (kind == HIPCLAW_SRC2_AXISYMMETRIC && (b->phys.maux < 1 || b->phys.meqn < 4))
        return set_error(HIPCLAW_ERR_ARG, "axisymmetric source needs meqn >= 4 and the radius in aux field 1");
    if (!b->staged.empty()) return set_error(HIPCLAW_ERR_STATE, "batch holds staged patches");
    b->src2 = kind;
    return HIPCLAW_OK;
}

int hipclaw_single_step_update(hipclaw_batch* b, double* q, const double* aux, double dx, double dy,
                               double t, double dt, int iter, int total_count, double* maxcfl)
{
    if (!b || !q || !maxcfl) return set_error(HIPCLAW_ERR_ARG, "NULL argument");
    if (total_count < 1 || iter < 0 || iter >= total_count)
        return set_error(HIPCLAW_ERR_ARG, "iter must lie in [0, total_count)");
    if (b->phys.maux > 0 && !aux) return set_error(HIPCLAW_ERR_ARG, "aux is NULL but maux > 0");
    if (!(dx > 0) || !(dy > 0) || !(dt >= 0)) return set_error(HIPCLAW_ERR_ARG, "bad dx, dy or dt");
    *maxcfl = 0.0;
    if (iter == 0 && !b->staged.empty()) b->staged.clear();  // an abandoned level
    if (b->have_dt && (dt != b->dt || t != b->t))
        return set_error(HIPCLAW_ERR_STATE, "t / dt changed inside one level update");
    b->t = t;
    b->dt = dt;
    b->have_dt = true;
    const int k = (int)b->staged.size();
    const long long per = b->T * b->phys.meqn, aper = b->T * b->phys.maux;
    const bool c5 = is_claw5(b->phys.solver);
    if (b->esz == 8 && !c5) {
        std::memcpy(static_cast<double*>(b->h_q) + k * per, q, sizeof(double) * per);
        if (aper) std::memcpy(static_cast<double*>(b->h_aux) + k * aper, aux, sizeof(double) * aper);
    } else if (b->esz == 8) {
        patch_from_host(q, static_cast<double*>(b->h_q) + k * per, b->phys.meqn, b->T, true);
        if (aper) patch_from_host(aux, static_cast<double*>(b->h_aux) + k * aper, b->phys.maux, b->T, true);
    } else {
        patch_from_host(q, static_cast<float*>(b->h_q) + k * per, b->phys.meqn, b->T, c5);
        if (aper) patch_from_host(aux, static_cast<float*>(b->h_aux) + k * aper, b->phys.maux, b->T, c5);
    }
    b->h_dxdy[2 * k] = dx;
    b->h_dxdy[2 * k + 1] = dy;
    b->staged.push_back({q});
    if ((int)b->staged.size() == b->cap || iter == total_count - 1) return flush(b, maxcfl);
    return HIPCLAW_OK;
}

}  // extern "C"
