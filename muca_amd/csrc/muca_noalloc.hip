// muca_noalloc.hip -- muca_reweight_noalloc: the reweighting of
// muca_reweight (include/muca.h) without any device allocation, for callers
// that cannot allocate (code under graph capture, tight serving loops).  Its
// scratch -- FP64 per-D moments, the crystal fields, the output planes and the
// range flag -- is a file-scope __device__ array of fixed size instead of
// buffers the context allocates on first use.
#include "ctx.h"

#include <mutex>

// One copy per GPU, resident from the library's load on that GPU to the end
// of the process.  Host code never passes `scratch` itself to a kernel or a
// copy (on the host the name only stands in for the symbol); it uses the
// address hipGetSymbolAddress returns for the current device.
static constexpr uint64_t kNoallocScratchDoubles = 4ull << 20;   // 32 MiB
static __device__ __align__(256) double scratch[kNoallocScratchDoubles];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_mu;
double *g_ptr[kMaxDevices] = {};

// The device address of `scratch` on `device`, looked up once per device
// (with that device current) on the first call there.
int scratch_of(int device, double **out)
{
    if (device < 0 || device >= kMaxDevices)
        return fail(MUCA_E_ARG, "device ordinal beyond the no-alloc scratch table");
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_ptr[device]) {
        HIP_TRY(hipSetDevice(device));
        void *p = nullptr;
        HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return fail(MUCA_E_HIP, "hipGetSymbolAddress(scratch) returned NULL");
        g_ptr[device] = (double *)p;
    }
    *out = g_ptr[device];
    return MUCA_OK;
}

}  // namespace

extern "C" {

int muca_reweight_noalloc(muca_ctx *c, const double *deltas, uint32_t ndelta, double *absM,
                          double *chi, double *binder)
{
    if (!c || !deltas || ndelta == 0) return fail(MUCA_E_ARG, "bad arguments");
    if (!c->joint && !c->moms)
        return fail(MUCA_E_STATE, "context has no joint histogram or per-D moments");
    if (!c->have_weights) return fail(MUCA_E_STATE, "muca_set_log_weight was not called");
    const uint64_t B = c->p.joint_blocks, nb = c->N + 1;
    // scratch layout (doubles): [0, 64) range flag; deltas; three output
    // planes of ndelta * B; FP64 moments [B][4][N+1] -- each 256-byte aligned
    auto al = [](uint64_t n) { return (n + 31) & ~(uint64_t)31; };
    const uint64_t off_deltas = 64, off_out = off_deltas + al(ndelta);
    const uint64_t off_mom = off_out + al(3ull * ndelta * B);
    const uint64_t need = off_mom + 4ull * B * nb;
    if (need > kNoallocScratchDoubles)
        return fail(MUCA_E_RANGE, "muca_reweight_noalloc needs " + std::to_string(need * 8) +
                                      " bytes of scratch, more than its " +
                                      std::to_string(kNoallocScratchDoubles * 8) +
                                      "-byte maximum (4 * blocks * (N+1) + 3 * ndelta * "
                                      "blocks + ndelta doubles); use muca_reweight");
    double *sc = nullptr;
    int rc = scratch_of(c->p.device, &sc);
    if (rc) return rc;
    HIP_TRY(hipSetDevice(c->p.device));
    uint32_t *flag = (uint32_t *)sc;
    double *d_deltas = sc + off_deltas, *out = sc + off_out, *mom = sc + off_mom;
    HIP_TRY(hipMemcpyAsync(d_deltas, deltas, ndelta * 8, hipMemcpyHostToDevice, c->st));
    if (c->joint) {
        HIP_TRY(muca::launch_joint_moments(c->st, c->joint, c->N, c->joint_stride, (uint32_t)B,
                                           mom));
    } else {
        HIP_TRY(muca::launch_fill_u32(c->st, flag, 1, 0));
        HIP_TRY(muca::launch_moments_fp64(c->st, c->moms, c->N, (uint32_t)B, mom, flag));
    }
    HIP_TRY(muca::launch_reweight(c->st, mom, c->lnw64, c->N, (uint32_t)B, c->p.beta, d_deltas,
                                  ndelta, out));
    const uint64_t plane = (uint64_t)ndelta * B;
    double *dst[3] = {absM, chi, binder};
    for (int k = 0; k < 3; k++)
        if (dst[k])
            HIP_TRY(hipMemcpyAsync(dst[k], out + k * plane, plane * 8, hipMemcpyDeviceToHost,
                                   c->st));
    uint32_t range = 0;
    if (!c->joint) HIP_TRY(hipMemcpyAsync(&range, flag, 4, hipMemcpyDeviceToHost, c->st));
    HIP_TRY(hipStreamSynchronize(c->st));
    if (range) return fail(MUCA_E_RANGE, "a D bin holds 2^32 or more entries");
    return MUCA_OK;
}

}  // extern "C"
