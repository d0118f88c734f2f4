// Allocation-free two-site H_eff matvec: ct_heff_apply with every
// intermediate (fused MPO coefficients, their padded copy, FP32 copies of the
// environments, T1 = L^T theta, the mixing output U and the split-K partial
// sums) placed in one fixed device array instead of the handle's growable
// workspaces, for callers that cannot call hipMalloc (code under HIP graph
// capture, tight serving loops).
//
// The array lives for the life of the process on every device the library
// runs on.  Its device address is resolved once per device when a handle is
// created on that device (hipGetSymbolAddress with the device current); host
// code only ever passes that resolved address, never the array's name.
#include <algorithm>
#include <mutex>

#include "ct_internal.h"

namespace ct {

namespace {

constexpr size_t kScratchBytes = (size_t)CT_NOALLOC_SCRATCH_BYTES;
constexpr int kMaxDevices = 64;

__device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

void* scratch_ptr(int device) {
    if (device < 0 || device >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    return g_scratch_ptr[device];
}

// Clears the handle's arena on every return path.
struct ArenaScope {
    Handle* h;
    ArenaScope(Handle* hh, Arena* a) : h(hh) { h->arena = a; }
    ~ArenaScope() { h->arena = nullptr; }
};

}  // namespace

int noalloc_scratch_init(int device) {
    if (device < 0 || device >= kMaxDevices) {
        set_last_error("device index beyond the allocation-free scratch table");
        return CT_STATUS_NOT_SUPPORTED;
    }
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[device]) return 0;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || p == nullptr) {
        set_last_error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        return e != hipSuccess ? CT_STATUS_HIP_ERROR_BASE + (int)e : CT_STATUS_NOT_SUPPORTED;
    }
    g_scratch_ptr[device] = p;
    return 0;
}

}  // namespace ct

extern "C" int ct_heff_apply_noalloc(ct_handle_t h, int prec, int dtype, const void* L, const void* Wi,
                                     const void* Wj, const void* R, const void* theta, void* out, int64_t Dl,
                                     int64_t Dlb, int64_t d, int64_t Dr, int64_t Drb, int64_t wL, int64_t wM,
                                     int64_t wR) {
    using namespace ct;
    CT_ARG(h != nullptr, 1, "handle is NULL");
    CT_ARG(prec == CT_PREC_FP64 || prec == CT_PREC_FP32, 2,
           "the allocation-free H_eff supports CT_PREC_FP64 and CT_PREC_FP32");
    CT_ARG(dtype == CT_F64 || dtype == CT_F32, 3, "unknown dtype");
    CT_ARG(prec != CT_PREC_FP64 || dtype == CT_F64, 3, "FP64 contraction requires FP64 buffers");
    CT_ARG(Dl > 0 && Dlb > 0 && d > 0 && Dr > 0 && Drb > 0 && wL > 0 && wM > 0 && wR > 0, 11,
           "dimensions must be > 0");
    const int64_t K = wL * d * d, NO = d * d * wR;
    CT_ARG(K <= 64 && NO <= 64, 16, "MPO bonds with w*d*d > 64 use the mixing GEMM path; call ct_heff_apply");
    CT_ARG(Dlb * d * d * Drb < (1ll << 31) && Dl * d * d * Dr < (1ll << 31) && wR * Dr < (1ll << 31), 11,
           "dimensions out of range");
    void* base = scratch_ptr(h->device);
    if (!base) {
        set_last_error("allocation-free scratch not resolved on this device");
        return CT_STATUS_NOT_SUPPORTED;
    }

    // Plan every workspace request of the call before anything is launched.
    const HeffDims hd{Dl, Dlb, d, Dr, Drb, wL, wM, wR};
    const int tAcc = prec == CT_PREC_FP64 ? CT_F64 : CT_F32;
    const size_t es = prec == CT_PREC_FP64 ? 8 : 4;
    const int64_t d2 = d * d;
    const int64_t nT1 = wL * Dlb * d2 * Dr, nU = Dlb * d2 * wR * Dr;
    const size_t sk = std::max(gemm_splitk_bytes(prec, heff_step1_desc(hd)), gemm_splitk_bytes(prec, heff_step4_desc(hd)));
    Arena a;
    a.base = (char*)base;
    a.cap = kScratchBytes;
    bool fits = arena_reserve(a, WS_W2, (size_t)(wL * d2 * d2 * wR) * 8) &&
                arena_reserve(a, tAcc == CT_F64 ? WS_W2PAD64 : WS_W2PAD32, mix_pad_bytes((int)K, (int)NO, tAcc));
    if (fits && prec == CT_PREC_FP32 && dtype == CT_F64)
        fits = arena_reserve(a, WS_LSPLIT, (size_t)(wL * Dl * Dlb) * 4) &&
               arena_reserve(a, WS_RSPLIT, (size_t)(wR * Dr * Drb) * 4) &&
               arena_reserve(a, WS_VSPLIT, (size_t)(Dl * d2 * Dr) * 4);
    if (fits) fits = arena_reserve(a, WS_T1, (size_t)nT1 * es) && arena_reserve(a, WS_T2, (size_t)nU * es);
    if (fits && sk > 0) fits = arena_reserve(a, WS_SPLITK, sk);
    if (!fits) {
        const size_t need = (size_t)(wL * d2 * d2 * wR) * 8 + (size_t)(nT1 + nU) * es + sk;
        set_last_error("ct_heff_apply_noalloc: these dimensions need about " + std::to_string(need >> 20) +
                       " MiB of scratch; the allocation-free limit is " + std::to_string(kScratchBytes >> 20) +
                       " MiB (use ct_heff_apply)");
        return CT_STATUS_SCRATCH_LIMIT;
    }

    ArenaScope scope(h, &a);
    void* w2 = nullptr;
    CT_CHECK(ws_get(h, WS_W2, (size_t)(wL * d2 * d2 * wR) * 8, &w2));
    CT_CHECK(heff_build_w2(h, dtype, Wi, Wj, wL, wM, wR, d, (double*)w2));
    HeffEnv env;
    CT_CHECK(heff_prepare(h, prec, dtype, L, R, (const double*)w2, hd, &env));
    return heff_apply(h, env, dtype, theta, out, nullptr);
}
